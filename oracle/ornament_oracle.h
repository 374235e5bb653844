/*
 * ornament_oracle.h -- CPU oracle for the zig_ornament path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker or as the timed CPU baseline, never as the thing measured or shipped.
 *
 * It restates, in plain C (FP32, no FMA contraction except where stated):
 *   - the WGSL megakernel  src/wgpu_backend/shaders/ (all ten .wgsl files)
 *   - the host iteration protocol  src/wgpu_backend/path_tracer.zig:212-265
 *   - the RNG seeding  src/wgpu_backend/buffers.zig:25-30
 *   - the BVH build  src/bvh.zig:33-383  (with Zig 0.11 std Xoshiro256 and
 *     std.sort.heap restated; exact node order is "parity unpinned")
 *   - the scene factories  src/scene.zig:64-351, src/aabb.zig:17-45,
 *     src/math.zig:13-47, src/camera.zig:24-68 and the zmath operations
 *     they call (zmath is an un-vendored submodule: "parity unpinned")
 * plus an FP64 brute-force closest-hit oracle that does not depend on the BVH.
 *
 * Binary layouts equal the reference's extern structs (gpu_structs.zig:22-117).
 */
#ifndef ORNAMENT_ORACLE_H
#define ORNAMENT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* gpu_structs.zig:22-31 (64 bytes) */
typedef struct orc_bvh_node {
    float lmin_or_v0[3];
    uint32_t left_or_custom_id;
    float lmax_or_v1[3];
    uint32_t right_or_material_index;
    float rmin_or_v2[3];
    uint32_t node_type; /* 0 internal, 1 sphere, 2 mesh, 3 triangle */
    float rmax_or_v3[3];
    uint32_t transform_id;
} orc_bvh_node;

/* gpu_structs.zig:33-59 (32 bytes) */
typedef struct orc_material {
    float albedo[3];
    uint32_t albedo_texture_index; /* 0xFFFFFFFF: no texture */
    float fuzz;
    float ior;
    uint32_t type; /* 0 lambertian, 1 metal, 2 dielectric, 3 diffuse light */
    uint32_t pad;
} orc_material;

/* gpu_structs.zig:88-117 (112 bytes) */
typedef struct orc_camera {
    float origin[3];
    float lens_radius;
    float lower_left_corner[3];
    uint32_t pad0;
    float horizontal[3];
    uint32_t pad1;
    float vertical[3];
    uint32_t pad2;
    float u[3];
    uint32_t pad3;
    float v[3];
    uint32_t pad4;
    float w[3];
    uint32_t pad5;
} orc_camera;

/* gpu_structs.zig:61-86 (144 bytes) */
typedef struct orc_constant_params {
    orc_camera camera;
    uint32_t depth;
    uint32_t width;
    uint32_t height;
    uint32_t flip_y;
    float inverted_gamma;
    float ray_cast_epsilon;
    uint32_t textures_count;
    float current_iteration;
} orc_constant_params;

/* One uploaded texture (wgpu_backend/buffers.zig:266-302). Formats: 8-bit
 * unorm with 1, 2 or 4 components, or 16-bit float (is_hdr) with 1, 2 or 4. */
typedef struct orc_texture {
    uint32_t width;
    uint32_t height;
    uint32_t num_components;
    uint32_t bytes_per_component;
    uint32_t is_hdr;
    uint32_t pad;
    const uint8_t *data;
} orc_texture;

/* The bound buffers of pathtracer.wgsl:7-17. */
typedef struct orc_scene {
    const orc_bvh_node *tlas;
    uint32_t tlas_len;
    const orc_bvh_node *blas;
    uint32_t blas_len;
    const float *normals; /* 4 floats per normal */
    uint32_t normals_len;
    const uint32_t *normal_indices;
    uint32_t normal_indices_len;
    const float *uvs; /* 2 floats per uv */
    uint32_t uvs_len;
    const uint32_t *uv_indices;
    uint32_t uv_indices_len;
    const float *transforms; /* 16 floats per matrix, WGSL column-major read */
    uint32_t transforms_len;
    const orc_material *materials;
    uint32_t materials_len;
    const orc_texture *textures;
    uint32_t textures_len;
} orc_scene;

/* Closest hit as returned by bvh_hit (bvh.wgsl:14-21). */
typedef struct orc_hit {
    uint32_t hit;
    float t;
    uint32_t material_index;
    uint32_t node_type;
    uint32_t inverted_transform_id;
    uint32_t tri_id;
    float uv[2];
} orc_hit;

/* Algorithmic traffic counters (instrumented oracle). */
typedef struct orc_stats {
    uint64_t rays;            /* bvh_hit calls */
    uint64_t node_visits;     /* BvhNode fetches, TLAS + BLAS */
    uint64_t transform_fetches; /* sphere tests + mesh entries */
    uint64_t sphere_hits;     /* shaded sphere hits */
    uint64_t mesh_hits;       /* shaded mesh hits */
    uint64_t textured_hits;   /* texel fetches */
    uint64_t samples;         /* camera samples */
} orc_stats;

/* ---------------- RNG (random.wgsl:11-22) ---------------- */
uint32_t orc_random_u32(uint32_t *state);
float orc_random_f32(uint32_t *state);
void orc_random_stream(uint32_t state, uint32_t n, uint32_t *out_u32, float *out_f32);
void orc_random_in_unit_sphere(uint32_t *state, float out[3]);
void orc_random_unit_vector(uint32_t *state, float out[3]);
void orc_random_in_unit_disk(uint32_t *state, float out[3]);

/* ------------- FP32 transcendental functions -------------
 * WGSL leaves the precision of pow/sin/cos/acos/atan2 to the implementation;
 * these are the oracle's definitions (see the .c header for the algorithms). */
float orc_sinf(float x);
float orc_cosf(float x);
float orc_powf(float x, float y);
float orc_acosf(float x);
float orc_atan2f(float y, float x);
/* Evaluate one of the above over arrays: op 0 sin, 1 cos, 2 pow, 3 acos, 4 atan2 */
void orc_eval_math(int op, const float *a, const float *b, float *out, uint32_t n);

/* ------------- camera (camera.wgsl:12-23, camera.zig:24-68) ------------- */
void orc_camera_init(const float lookfrom[3], const float lookat[3], const float vup[3],
                     float aspect_ratio, float vfov, float aperture, float focus_dist,
                     orc_camera *out);
void orc_camera_get_ray(const orc_camera *cam, float s, float t, uint32_t *rng,
                        float origin[3], float direction[3]);

/* ------------- intersection (bvh.wgsl) ------------- */
int orc_bvh_hit(const orc_scene *sc, float t_min, const float origin[3],
                const float direction[3], orc_hit *out, orc_stats *stats);
void orc_trace_rays(const orc_scene *sc, float t_min, const float *rays /* 6 per ray */,
                    uint32_t n, orc_hit *out, int nthreads);
/* Single-primitive tests, exposed for closed-form tests. */
void orc_safe_invdir(const float d[3], float out[3]);
void orc_aabb_hit(const float mn[3], const float mx[3], const float invdir[3],
                  const float oxinvdir[3], float t_min, float t_max, float out[2]);
float orc_triangle_hit(const float o[3], const float d[3], const float v0[3],
                       const float v1[3], const float v2[3], float t_min, float t_max,
                       float uv[2]);
float orc_sphere_hit(const float o[3], const float d[3], float t_min, float t_max);

/* FP64 brute force over every primitive, independent of the BVH.
 * out8 per ray: [hit, t, transform_id, tri (global triangle index or
 * 0xFFFFFFFF), ambiguous flag, second-best t, conditioning, 0].
 * A ray is ambiguous when a primitive lies within rel_eps of an edge, of
 * t_min or of tangency at a distance that could beat the best hit, or when
 * the two best hits are within rel_eps of each other (ties).
 * conditioning: first-order forward error bound, relative to t, of the
 * reference's FP32 transform + intersection sequence for the best hit. */
void orc_brute_force(const orc_scene *sc, double t_min, const float *rays, uint32_t n,
                     double rel_eps, double *out8, int nthreads);

/* ------------- shading + whole image (pathtracer.wgsl) ------------- */
/* One render() call of path_tracer.zig:252-265 on every pixel. `params`
 * holds current_iteration BEFORE this call (State.current_iteration); the
 * function advances it like update() does. acc: 4*W*H floats, rng: W*H,
 * fb: 4*W*H floats. Returns the new current_iteration. */
float orc_render(const orc_scene *sc, const orc_constant_params *params, uint32_t iterations,
                 float *acc, uint32_t *rng, float *fb, int nthreads, orc_stats *stats);
/* Same on a subset of pixels: acc/rng/fb hold one entry per listed pixel. */
float orc_render_pixels(const orc_scene *sc, const orc_constant_params *params,
                        uint32_t iterations, const uint32_t *pixels, uint32_t n, float *acc,
                        uint32_t *rng, float *fb, int nthreads, orc_stats *stats);
/* The pieces of ray_color, exposed for closed-form tests; ray_color runs
 * exactly these functions. */
typedef struct orc_hit_record {  /* hitrecord.wgsl:1-8 */
    float p[3];
    float normal[3];  /* shading normal, flipped to face the ray */
    float uv[2];
    uint32_t material_index;
    float t;
    int32_t front_face;
} orc_hit_record;
/* The hit record of a closest hit (pathtracer.wgsl:111-144 +
 * hitrecord.wgsl:10-18): sphere normal from the world-space centre and uv
 * from acos/atan2; mesh normal and uv interpolated and the normal taken
 * through transpose(inverse) (transform.wgsl:16-19). */
void orc_make_hit_record(const orc_scene *sc, const float ro[3], const float rd[3], const orc_hit *hit,
                         orc_hit_record *out);
/* material_scatter (material.wgsl:61-77): returns 1 when the path goes on. */
int orc_material_scatter(const orc_scene *sc, const orc_constant_params *p, const float rd[3],
                         const orc_hit_record *h, uint32_t *rng, float att[3], float origin[3],
                         float direction[3]);
/* material_emit (material.wgsl:79-89). */
void orc_material_emit(const orc_scene *sc, const orc_constant_params *p, const orc_hit_record *h,
                       float out[3]);
/* textureSampleLevel(t, default sampler, uv, 0) (material.wgsl:91-99,
 * webgpu.zig:812-825). */
void orc_sample_texture(const orc_texture *tx, const float uv[2], float out[3]);
float orc_half_to_float(uint16_t h);

/* RNG buffer seeding (buffers.zig:25-30): state = index + seed * pixel_count. */
void orc_seed_rng(uint32_t *rng, uint32_t pixel_count, uint32_t seed);

/* ------------- host side: zmath, scene factories, BVH build ------------- */
/* Row-vector 4x4 matrices stored as 16 floats, rows contiguous (zmath.Mat). */
void orc_mat_identity(float m[16]);
void orc_mat_mul(const float a[16], const float b[16], float out[16]);
void orc_mat_scaling(float sx, float sy, float sz, float out[16]);
void orc_mat_translation(float x, float y, float z, float out[16]);
void orc_mat_from_quat(const float q[4], float out[16]);
int orc_mat_inverse(const float m[16], float out[16]);
void orc_vec_mul_mat(const float v[4], const float m[16], float out[4]);
void orc_rotation_between_vectors(const float a[4], const float b[4], float out[16]);
void orc_aabb_transform(const float m[16], const float mn[4], const float mx[4],
                        float omn[4], float omx[4]);

/* BVH build input: the scene's object lists, in creation order
 * (scene.zig:17-21). Matrices are zmath row-major. */
typedef struct orc_build_material {
    float albedo[4];
    int32_t texture; /* scene texture index, or -1 for a plain colour */
    float fuzz;
    float ior;
    uint32_t type;
} orc_build_material;

typedef struct orc_build_sphere {
    float transform[16];
    float aabb_min[4];
    float aabb_max[4];
    uint32_t material;
} orc_build_sphere;

typedef struct orc_build_mesh {
    const float *vertices; /* 4 per vertex */
    uint32_t vertices_len;
    const uint32_t *vertex_indices;
    uint32_t vertex_indices_len;
    const float *normals; /* 4 per normal */
    uint32_t normals_len;
    const uint32_t *normal_indices;
    uint32_t normal_indices_len;
    const float *uvs; /* 2 per uv */
    uint32_t uvs_len;
    const uint32_t *uv_indices;
    uint32_t uv_indices_len;
    float transform[16];
    float aabb_min[4];
    float aabb_max[4];
    uint32_t material;
} orc_build_mesh;

typedef struct orc_build_instance {
    uint32_t mesh;
    float transform[16];
    float aabb_min[4];
    float aabb_max[4];
    uint32_t material;
} orc_build_instance;

typedef struct orc_build_input {
    const orc_build_material *materials;
    uint32_t materials_len;
    const orc_build_sphere *spheres;
    uint32_t spheres_len;
    const orc_build_mesh *meshes;
    uint32_t meshes_len;
    const orc_build_instance *instances;
    uint32_t instances_len;
} orc_build_input;

/* Output arrays are allocated by the oracle; free with orc_bvh_free. */
typedef struct orc_build_output {
    orc_bvh_node *tlas;
    uint32_t tlas_len;
    orc_bvh_node *blas;
    uint32_t blas_len;
    float *normals;
    uint32_t normals_len;
    uint32_t *normal_indices;
    uint32_t normal_indices_len;
    float *uvs;
    uint32_t uvs_len;
    uint32_t *uv_indices;
    uint32_t uv_indices_len;
    float *transforms;
    uint32_t transforms_len;
    orc_material *materials;
    uint32_t materials_len;
    uint32_t *textures; /* scene texture index of each bound texture */
    uint32_t textures_len;
} orc_build_output;

int orc_bvh_build(const orc_build_input *in, orc_build_output *out);
/* The same build drawing its split axes from the caller's Xoshiro256 state,
 * which it leaves advanced: the reference's one generator per process
 * (bvh.zig:12-13) across consecutive builds. orc_bvh_build is this with a
 * state freshly seeded with 1244. */
int orc_bvh_build_stream(const orc_build_input *in, orc_build_output *out, uint64_t prng[4]);
void orc_bvh_free(orc_build_output *out);

/* Zig 0.11 std.rand.Xoshiro256 seeded through SplitMix64, and
 * Random.intRangeAtMost(usize, 0, 2); exposed for known-answer tests. */
uint64_t orc_splitmix64_next(uint64_t *state);
void orc_xoshiro_seed(uint64_t seed, uint64_t s[4]);
uint64_t orc_xoshiro_next(uint64_t s[4]);
uint64_t orc_xoshiro_range_at_most(uint64_t s[4], uint64_t at_most);
/* std.sort.heap over floats with the `<` comparator (ties are order-relevant
 * for the BVH); sorts `keys` in place and permutes `ids` alongside. */
void orc_heap_sort_keys(float *keys, uint32_t *ids, uint32_t n);

/* Scene factories (scene.zig). Meshes: vertices/normals are 4 floats each. */
void orc_create_sphere(const float center[4], float radius, float transform[16],
                       float aabb_min[4], float aabb_max[4]);
void orc_plane_mesh_transform(const float center[4], float side1, float side2,
                              const float normal[4], float transform[16]);
void orc_mesh_aabb(const float *vertices, const uint32_t *vertex_indices,
                   uint32_t vertex_indices_len, const float transform[16],
                   float not_transformed_min[4], float not_transformed_max[4],
                   float aabb_min[4], float aabb_max[4]);

#ifdef __cplusplus
}
#endif

#endif
