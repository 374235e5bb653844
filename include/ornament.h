/*
 * ornament.h -- C-ABI of the ornament host layer, restated in C++ for MI355X.
 *
 * It mirrors the public API of src/ornament.zig: Scene and its factories
 * (scene.zig), Camera (camera.zig), State (state.zig), the BVH build
 * (bvh.zig) and the duck-typed PathTracer interface {init, state, scene,
 * setResolution, render, getFrameBuffer, deinit} that the reference's
 * backends implement (wgpu_backend/path_tracer.zig:14-266,
 * consumed by glfw_example/app_no_interop.zig:50-185). The PathTracer here
 * drives the HIP kernels through ornament_hip.h.
 *
 * Vectors are zmath.Vec-shaped (4 floats: x, y, z, w) and matrices are
 * zmath.Mat-shaped (16 floats, row-vector convention, rows contiguous).
 * Objects returned by the scene factories are owned by their scene, as in
 * the reference (scene.zig:369-379); a PathTracer takes ownership of its
 * scene and frees it in deinit (wgpu_backend/path_tracer.zig:127).
 */
#ifndef ORNAMENT_H
#define ORNAMENT_H

#include <stddef.h>
#include <stdint.h>

#include "ornament_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ornament_scene ornament_scene;
typedef struct ornament_material ornament_material;
typedef struct ornament_texture ornament_texture;
typedef struct ornament_sphere ornament_sphere;
typedef struct ornament_mesh ornament_mesh;
typedef struct ornament_mesh_instance ornament_mesh_instance;
typedef struct ornament_bvh ornament_bvh;
typedef struct ornament_path_tracer ornament_path_tracer;

/* ---------------- Scene (scene.zig:13-380) ---------------- */
orn_status ornament_scene_init(ornament_scene **out);                         /* scene.zig:27 */
void ornament_scene_deinit(ornament_scene *scene);                            /* scene.zig:50 */

/* Materials. A Color is either a vector or a texture (color.zig:4-7). */
orn_status ornament_scene_lambertian(ornament_scene *s, const float albedo[4],
                                     ornament_material **out);                /* scene.zig:64 */
orn_status ornament_scene_lambertian_texture(ornament_scene *s, ornament_texture *albedo,
                                             ornament_material **out);
orn_status ornament_scene_metal(ornament_scene *s, const float albedo[4], float fuzz,
                                ornament_material **out);                     /* scene.zig:78 */
orn_status ornament_scene_metal_texture(ornament_scene *s, ornament_texture *albedo, float fuzz,
                                        ornament_material **out);
orn_status ornament_scene_dielectric(ornament_scene *s, float ior,
                                     ornament_material **out);                /* scene.zig:92 */
orn_status ornament_scene_diffuse_light(ornament_scene *s, const float albedo[4],
                                        ornament_material **out);             /* scene.zig:106 */
orn_status ornament_scene_diffuse_light_texture(ornament_scene *s, ornament_texture *albedo,
                                                ornament_material **out);

/* Shapes */
orn_status ornament_scene_create_sphere(ornament_scene *s, const float center[4], float radius,
                                        ornament_material *m,
                                        ornament_sphere **out);               /* scene.zig:124 */
orn_status ornament_scene_create_sphere_mesh(ornament_scene *s, const float center[4],
                                             float radius, ornament_material *m,
                                             ornament_mesh **out);            /* scene.zig:140 */
orn_status ornament_scene_create_plane_mesh(ornament_scene *s, const float center[4],
                                            float side1_length, float side2_length,
                                            const float normal[4], ornament_material *m,
                                            ornament_mesh **out);             /* scene.zig:235 */
/* scene.zig:263. vertices and normals: 4 floats each; uvs: 2 floats each.
 * With no uvs every vertex gets (0.5, 0.5) (scene.zig:303-305). */
orn_status ornament_scene_create_mesh(ornament_scene *s, const float *vertices, size_t vertices_len,
                                      const uint32_t *vertex_indices, size_t vertex_indices_len,
                                      const float *normals, size_t normals_len,
                                      const uint32_t *normal_indices, size_t normal_indices_len,
                                      const float *uvs, size_t uvs_len, const uint32_t *uv_indices,
                                      size_t uv_indices_len, const float transform[16],
                                      ornament_material *m, ornament_mesh **out);
orn_status ornament_scene_create_mesh_instance(ornament_scene *s, ornament_mesh *mesh,
                                               const float transform[16], ornament_material *m,
                                               ornament_mesh_instance **out); /* scene.zig:323 */
/* scene.zig:339; data holds width*height*num_components*bytes_per_component bytes */
orn_status ornament_scene_create_texture(ornament_scene *s, const uint8_t *data, size_t data_len,
                                         uint32_t width, uint32_t height, uint32_t num_components,
                                         uint32_t bytes_per_component, int is_hdr, float gamma,
                                         ornament_texture **out);
orn_status ornament_scene_attach_sphere(ornament_scene *s, ornament_sphere *x);     /* :357 */
orn_status ornament_scene_attach_mesh(ornament_scene *s, ornament_mesh *x);         /* :361 */
orn_status ornament_scene_attach_mesh_instance(ornament_scene *s, ornament_mesh_instance *x);

/* Object counts: spheres, meshes, mesh instances, materials, textures. */
orn_status ornament_scene_counts(const ornament_scene *s, uint32_t counts[5]);

/* ---------------- Camera (camera.zig:4-114) ---------------- */
/* Replaces scene.camera with Camera.init(...) (camera.zig:24-68). */
orn_status ornament_scene_set_camera(ornament_scene *s, const float lookfrom[4],
                                     const float lookat[4], const float vup[4], float aspect_ratio,
                                     float vfov, float aperture, float focus_dist);
orn_status ornament_camera_set_look_at(ornament_scene *s, const float lookfrom[4],
                                       const float lookat[4], const float vup[4]); /* :98 */
orn_status ornament_camera_set_aspect_ratio(ornament_scene *s, float aspect_ratio);  /* :70 */
/* The device view of the camera (gpu_structs.Camera.from, gpu_structs.zig:103-116). */
orn_status ornament_camera_get(const ornament_scene *s, orn_camera *out);
orn_status ornament_camera_get_dirty(const ornament_scene *s, int *dirty);

/* ---------------- Bvh (bvh.zig:15-109) ---------------- */
/* Builds the TLAS/BLAS and the attribute arrays from scene.spheres,
 * scene.meshes and scene.mesh_instances. Like the reference it records
 * material_index / texture_id / bvh_id on the scene's objects. */
orn_status ornament_bvh_init(ornament_scene *s, int row_major_transforms, ornament_bvh **out);
void ornament_bvh_deinit(ornament_bvh *bvh);
/* The generator of the builds' random split axes (bvh.zig:12-13: one
 * DefaultPrng per process, seeded with 1244, and bvh.zig:239, 331). The
 * reference shares it across every Bvh.init of the process, so a build
 * depends on the builds before it. mode 1 does the same: the builds that
 * follow continue one stream (one build at a time), which this call restarts
 * from seed 1244, as a new process would. mode 0 (the default) seeds each
 * build with 1244: every build is then the one the reference makes first in
 * a process. Other modes: InvalidValue. */
orn_status ornament_bvh_shared_generator(int mode);
/* Pointers into the Bvh's arrays (valid until ornament_bvh_deinit); textures
 * point into the scene's texture data. */
orn_status ornament_bvh_arrays(const ornament_bvh *bvh, orn_scene_desc *out);
/* Scene texture index behind each bound texture (bvh.textures). */
orn_status ornament_bvh_texture_sources(const ornament_bvh *bvh, uint32_t *dst, size_t len);

/* ---------------- PathTracer (the backend interface) ---------------- */
/* PathTracer.init(allocator, scene): takes ownership of the scene, builds the
 * BVH and uploads it. device < 0: the last HIP device. */
orn_status ornament_path_tracer_init(int device, ornament_scene *scene,
                                     ornament_path_tracer **out);
orn_status ornament_path_tracer_deinit(ornament_path_tracer *pt);
/* The public `scene` field: camera mutations go through the scene. */
ornament_scene *ornament_path_tracer_scene(ornament_path_tracer *pt);
orn_status ornament_path_tracer_set_resolution(ornament_path_tracer *pt, uint32_t width,
                                               uint32_t height);
orn_status ornament_path_tracer_render(ornament_path_tracer *pt);
/* dst holds `pixels` float4 entries (getFrameBuffer, path_tracer.zig:161). */
orn_status ornament_path_tracer_get_framebuffer(ornament_path_tracer *pt, float *dst,
                                                size_t pixels);
orn_status ornament_path_tracer_get_accumulation(ornament_path_tracer *pt, float *dst,
                                                 size_t pixels);
/* GPU milliseconds of the last render() (kernels only). */
orn_status ornament_path_tracer_last_render_ms(ornament_path_tracer *pt, float *ms);

/* ---------------- State (state.zig:3-80), the public `state` field ------- */
orn_status ornament_state_set_depth(ornament_path_tracer *pt, uint32_t depth);
orn_status ornament_state_get_depth(ornament_path_tracer *pt, uint32_t *depth);
orn_status ornament_state_set_iterations(ornament_path_tracer *pt, uint32_t iterations);
orn_status ornament_state_get_iterations(ornament_path_tracer *pt, uint32_t *iterations);
orn_status ornament_state_set_gamma(ornament_path_tracer *pt, float gamma);
orn_status ornament_state_get_gamma(ornament_path_tracer *pt, float *gamma);
orn_status ornament_state_set_flip_y(ornament_path_tracer *pt, int flip_y);
orn_status ornament_state_get_flip_y(ornament_path_tracer *pt, int *flip_y);
orn_status ornament_state_set_ray_cast_epsilon(ornament_path_tracer *pt, float epsilon);
orn_status ornament_state_get_ray_cast_epsilon(ornament_path_tracer *pt, float *epsilon);
orn_status ornament_state_get_resolution(ornament_path_tracer *pt, uint32_t *width,
                                         uint32_t *height);
orn_status ornament_state_get_current_iteration(ornament_path_tracer *pt, float *iteration);
/* ConstantParams.from(camera, state, textures_count) (gpu_structs.zig:73-85). */
orn_status ornament_path_tracer_constant_params(ornament_path_tracer *pt,
                                                orn_constant_params *out);

/* ------------------------------------------------------------------ */
/* Scene files and images. The reference builds its scenes in code     */
/* (glfw_example/examples.zig) and shows images in a window; these     */
/* files make a scene and its render settings reproducible between      */
/* machines and keep rendered images (Target.getFrameBuffer,            */
/* wgpu_backend/buffers.zig:67-91) as float data.                       */
/* ------------------------------------------------------------------ */

/* Render settings stored beside a scene (State, state.zig:3-23). */
typedef struct ornament_render_settings {
    uint32_t width;
    uint32_t height;
    uint32_t iterations; /* samples per pixel of one render() */
    uint32_t depth;
    float gamma;
    float ray_cast_epsilon;
    uint32_t flip_y;
    uint32_t seed; /* RNG seed of the target (orn_target_create) */
} ornament_render_settings;

/* Scene file (little-endian, "ORNSCENE" version 1): the camera, optional
 * render settings, then every texture, material, sphere, mesh and mesh
 * instance in creation order with its attach flag; meshes are stored as
 * created (vertices, normals and uvs with their index arrays, model
 * transform). Loading replays the scene factories, so a loaded scene builds
 * the same BVH as the saved one. `settings` may be NULL (not stored / not
 * wanted); `has_settings` reports whether the file held them. */
orn_status ornament_scene_save(const ornament_scene *s, const ornament_render_settings *settings,
                               const char *path);
orn_status ornament_scene_load(const char *path, ornament_scene **out,
                               ornament_render_settings *settings, int *has_settings);

/* Images: `rgba` holds width * height float4 pixels, row 0 first; row 0 is
 * the bottom of the picture (v = 0 in camera_get_ray, camera.wgsl:12-23),
 * as a framebuffer rendered with flip_y = false holds it.
 * PFM ("PF", RGB, little-endian, rows bottom to top; alpha is dropped). */
orn_status ornament_image_write_pfm(const char *path, const float *rgba, uint32_t width,
                                    uint32_t height);
/* Reads a PFM ("PF" RGB or "Pf" grey, either byte order) into rgba (alpha
 * 1); width and height are returned; `pixels` is the capacity of rgba (call
 * with rgba = NULL to query the size). */
orn_status ornament_image_read_pfm(const char *path, float *rgba, size_t pixels, uint32_t *width,
                                   uint32_t *height);
/* OpenEXR 2.0 scanline image, uncompressed, channels A, B, G, R as 32-bit
 * float, rows top to bottom. */
orn_status ornament_image_write_exr(const char *path, const float *rgba, uint32_t width,
                                    uint32_t height);

#ifdef __cplusplus
}
#endif

#endif
