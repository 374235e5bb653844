/*
 * ornament_oracle.c -- CPU oracle for the zig_ornament path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py as the checker / CPU baseline. The product
 * (zig_ornament_amd) never links, loads or calls it.
 *
 * Compile with -ffp-contract=off: every WGSL expression is evaluated as
 * written, one IEEE-754 rounding per operation, left to right. The one
 * deliberate fused multiply-add is the slab test of aabb_hit
 * (bvh.wgsl:185-186, `mmax * invdir + oxinvdir`), which the reference laid
 * out for FMA by precomputing oxinvdir = -origin * invdir (bvh.wgsl:36).
 *
 * Precision of WGSL builtins is left to the implementation by the WGSL
 * spec. The oracle defines them as:
 *   sqrt, division    IEEE-754 correctly rounded
 *   normalize(v)      v * (1 / sqrt(dot(v, v)))
 *   sin, cos          Cody-Waite reduction by pi/2 (2-part constant, FMA)
 *                     and Taylor polynomials on [-pi/4, pi/4]
 *   pow(x, y)         exp2(y * log2(x)) as in the WGSL spec's error bound,
 *                     with log2 from the atanh series and exp2 from a
 *                     degree-7 polynomial in the fraction
 *   acos, atan2       Cephes asinf / atanf polynomials (S. L. Moshier)
 * The tests pin each against libm in double precision.
 *
 * Parity status of each part:
 *   RNG (random.wgsl)            pinned: closed form + golden streams
 *   camera (camera.wgsl)         pinned: closed-form invariants
 *   slab/triangle/sphere tests   pinned: closed forms + FP64 brute force
 *   bvh_hit                      pinned: FP64 brute force over primitives
 *   materials / ray_color        pinned: furnace tests, draw counts, closed forms
 *   dielectric / metal scatter   pinned: Snell and mirror laws, Schlick
 *                                decision vs the draw, TIR without a draw,
 *                                absorption below the surface
 *   texture fetch, half floats   pinned: exact nearest/clamp texel selection,
 *                                byte/255, all 65,536 binary16 encodings
 *   sphere uv, mesh normal       pinned: axis-point closed forms; inverse-
 *                                transpose normal of a non-uniform instance
 *   BVH build node order         parity unpinned (needs Zig 0.11 std to run);
 *                                invariants pinned
 *   zmath host arithmetic        parity unpinned (zmath is not vendored);
 *                                closed forms pinned
 */
#include "ornament_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    float x, y, z;
} v3;

static inline v3 mk(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 vadd(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 vscale(float s, v3 a) { return mk(s * a.x, s * a.y, s * a.z); }
static inline v3 vneg(v3 a) { return mk(-a.x, -a.y, -a.z); }
static inline float vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 vcross(v3 a, v3 b)
{
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline v3 vnormalize(v3 a)
{
    float inv = 1.0f / sqrtf(vdot(a, a));
    return mk(a.x * inv, a.y * inv, a.z * inv);
}
static inline v3 vload(const float *p) { return mk(p[0], p[1], p[2]); }
static inline void vstore(float *p, v3 a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }

/* IEEE-754 minNum / maxNum: a NaN operand yields the other operand (the
 * semantics of C fminf/fmaxf), written inline so the compiler can use
 * vector min/max instructions. The sign of a zero result is not specified
 * and never reaches a branch in this file. */
static inline float ominf(float a, float b) { return (a < b) ? a : ((b != b) ? a : b); }
static inline float omaxf(float a, float b) { return (a > b) ? a : ((b != b) ? a : b); }

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* utility.wgsl:1 */
#define PI_F 3.1415927410125732f
#define PIO2_F 1.5707963705062866f
#define PIO4_F 0.7853981852531433f

/* ===================================================================== */
/* RNG: random.wgsl:11-87, seeded per pixel by buffers.zig:25-30          */
/* ===================================================================== */

uint32_t orc_random_u32(uint32_t *state)
{
    /* random.wgsl:14-17: an addition of two constants, not a PCG multiply */
    uint32_t old_state = *state + 747796405u + 2891336453u;
    uint32_t word = ((old_state >> ((old_state >> 28u) + 4u)) ^ old_state) * 277803737u;
    *state = (word >> 22u) ^ word;
    return *state;
}

float orc_random_f32(uint32_t *state)
{
    /* random.wgsl:21: f32(u32) rounds to nearest, so 1.0 is reachable */
    return (float)orc_random_u32(state) * (1.0f / 4294967296.0f);
}

void orc_random_stream(uint32_t state, uint32_t n, uint32_t *out_u32, float *out_f32)
{
    uint32_t s = state;
    for (uint32_t i = 0; i < n; i++) {
        uint32_t a = s;
        if (out_u32) out_u32[i] = orc_random_u32(&a);
        uint32_t b = s;
        if (out_f32) out_f32[i] = orc_random_f32(&b);
        orc_random_u32(&s);
    }
}

/* ===================================================================== */
/* FP32 transcendental functions (stand-ins for WGSL builtins)            */
/* ===================================================================== */

static void sincos_f(float x, float *s, float *c)
{
    /* k = round(x * 2/pi); r = x - k*pi/2 with pi/2 split in two floats */
    float kf = floorf(x * 0.6366197466850281f + 0.5f);
    float r = fmaf(kf, -1.5707963705062866f, x);
    r = fmaf(kf, 4.371138828673793e-08f, r);
    int q = ((int)kf) & 3;
    float r2 = r * r;
    /* sin r = r + r^3 (-1/6 + r^2 (1/120 + r^2 (-1/5040 + r^2/362880))) */
    float ps = fmaf(r2, 2.7557318844628753e-06f, -0.00019841270113829523f);
    ps = fmaf(r2, ps, 0.008333333767950535f);
    ps = fmaf(r2, ps, -0.1666666716337204f);
    float sp = fmaf(r * r2, ps, r);
    /* cos r = 1 + r^2 (-1/2 + r^2 (1/24 + r^2 (-1/720 + r^2 (1/40320 - r^2/3628800)))) */
    float pc = fmaf(r2, -2.755731998149713e-07f, 2.4801587642286904e-05f);
    pc = fmaf(r2, pc, -0.0013888889225199819f);
    pc = fmaf(r2, pc, 0.0416666679084301f);
    pc = fmaf(r2, pc, -0.5f);
    float cp = fmaf(r2, pc, 1.0f);
    switch (q) {
    case 0: *s = sp; *c = cp; break;
    case 1: *s = cp; *c = -sp; break;
    case 2: *s = -sp; *c = -cp; break;
    default: *s = -cp; *c = sp; break;
    }
}

float orc_sinf(float x)
{
    float s, c;
    sincos_f(x, &s, &c);
    return s;
}

float orc_cosf(float x)
{
    float s, c;
    sincos_f(x, &s, &c);
    return c;
}

/* log2 for finite x > 0 */
static float log2_pos(float x)
{
    uint32_t bits = f2u(x);
    int e = 0;
    if (bits < 0x00800000u) { /* subnormal */
        x = x * 16777216.0f;
        bits = f2u(x);
        e = -24;
    }
    e += (int)((bits >> 23) & 0xffu) - 127;
    float m = u2f((bits & 0x007fffffu) | 0x3f800000u);
    if (m > 1.4142135381698608f) {
        m = m * 0.5f;
        e += 1;
    }
    /* ln m = 2 atanh(s), s = (m-1)/(m+1), |s| <= 0.1716 */
    float s = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float p = fmaf(s2, 0.09090909361839294f, 0.1111111119389534f);
    p = fmaf(s2, p, 0.1428571492433548f);
    p = fmaf(s2, p, 0.20000000298023224f);
    p = fmaf(s2, p, 0.3333333432674408f);
    p = fmaf(s2, p, 1.0f);
    float ln_m = 2.0f * s * p;
    return fmaf(ln_m, 1.4426950216293335f, (float)e);
}

static float exp2_f(float z)
{
    if (z != z) return z;
    if (z >= 128.0f) return INFINITY;
    if (z < -150.0f) return 0.0f;
    float n = floorf(z + 0.5f);
    float f = z - n; /* exact, |f| <= 0.5 */
    /* 2^f = sum (f ln2)^k / k!, k <= 7 */
    float p = fmaf(f, 1.5252733646775596e-05f, 0.0001540352968731895f);
    p = fmaf(f, p, 0.0013333557872101665f);
    p = fmaf(f, p, 0.009618128649890423f);
    p = fmaf(f, p, 0.05550410971045494f);
    p = fmaf(f, p, 0.24022650718688965f);
    p = fmaf(f, p, 0.6931471824645996f);
    p = fmaf(f, p, 1.0f);
    int ni = (int)n;
    if (ni >= -126) {
        if (ni > 127) return INFINITY;
        return p * u2f((uint32_t)(ni + 127) << 23);
    }
    /* result lands in the subnormal range: scale in two exact steps */
    float t = p * u2f((uint32_t)(ni + 64 + 127) << 23);
    return t * u2f((uint32_t)(-64 + 127) << 23);
}

float orc_powf(float x, float y)
{
    if (x != x || y != y) return x + y;
    if (x < 0.0f) return NAN;
    if (x == 0.0f) return (y > 0.0f) ? 0.0f : ((y == 0.0f) ? 1.0f : INFINITY);
    if (x == INFINITY) return (y > 0.0f) ? INFINITY : ((y == 0.0f) ? 1.0f : 0.0f);
    return exp2_f(y * log2_pos(x));
}

/* Cephes asinf kernel on [0, 0.5]: t + t^3 P(t^2) */
static float asin_kernel(float t, float z)
{
    float p = fmaf(z, 4.2163199048e-2f, 2.4181311049e-2f);
    p = fmaf(z, p, 4.5470025998e-2f);
    p = fmaf(z, p, 7.4953002686e-2f);
    p = fmaf(z, p, 1.6666752422e-1f);
    return fmaf(p * z, t, t);
}

float orc_acosf(float x)
{
    float a = fabsf(x);
    if (a > 0.5f) {
        float z = 0.5f * (1.0f - a);
        float s = sqrtf(z);
        float r = 2.0f * asin_kernel(s, z);
        return (x < 0.0f) ? (PI_F - r) : r;
    }
    return PIO2_F - asin_kernel(x, x * x);
}

float orc_atan2f(float y, float x)
{
    if (x != x || y != y) return x + y;
    float ax = fabsf(x), ay = fabsf(y);
    float mx = ax > ay ? ax : ay;
    float mn = ax > ay ? ay : ax;
    float a;
    if (mx == 0.0f) {
        a = 0.0f;
    } else {
        float t = mn / mx;
        float base = 0.0f;
        if (t > 0.4142135679721832f) {
            base = PIO4_F;
            t = (t - 1.0f) / (t + 1.0f);
        }
        float z = t * t;
        /* Cephes atanf: t + t^3 Q(t^2) */
        float p = fmaf(z, 8.05374449538e-2f, -1.38776856032e-1f);
        p = fmaf(z, p, 1.99777106478e-1f);
        p = fmaf(z, p, -3.33329491539e-1f);
        a = base + fmaf(p * z, t, t);
    }
    if (ay > ax) a = PIO2_F - a;
    if (x < 0.0f || (x == 0.0f && signbit(x))) a = PI_F - a;
    if (signbit(y)) a = -a;
    return a;
}

void orc_eval_math(int op, const float *a, const float *b, float *out, uint32_t n)
{
    for (uint32_t i = 0; i < n; i++) {
        switch (op) {
        case 0: out[i] = orc_sinf(a[i]); break;
        case 1: out[i] = orc_cosf(a[i]); break;
        case 2: out[i] = orc_powf(a[i], b[i]); break;
        case 3: out[i] = orc_acosf(a[i]); break;
        default: out[i] = orc_atan2f(a[i], b[i]); break;
        }
    }
}

/* random.wgsl:36-53: theta is uniform in [0, pi] (not acos-distributed) */
static v3 random_in_unit_sphere(uint32_t *rng)
{
    float r = orc_powf(orc_random_f32(rng), 0.33333f);
    float theta = PI_F * orc_random_f32(rng);
    float phi = (2.0f * PI_F) * orc_random_f32(rng);
    float st, ct, sp, cp;
    sincos_f(theta, &st, &ct);
    sincos_f(phi, &sp, &cp);
    return mk(r * st * cp, r * st * sp, r * ct);
}

/* random.wgsl:55-57 */
static v3 random_unit_vector(uint32_t *rng) { return vnormalize(random_in_unit_sphere(rng)); }

/* random.wgsl:69-87 */
static v3 random_in_unit_disk(uint32_t *rng)
{
    float r = sqrtf(orc_random_f32(rng));
    float alpha = (2.0f * PI_F) * orc_random_f32(rng);
    float s, c;
    sincos_f(alpha, &s, &c);
    return mk(r * c, r * s, 0.0f);
}

void orc_random_in_unit_sphere(uint32_t *state, float out[3]) { vstore(out, random_in_unit_sphere(state)); }
void orc_random_unit_vector(uint32_t *state, float out[3]) { vstore(out, random_unit_vector(state)); }
void orc_random_in_unit_disk(uint32_t *state, float out[3]) { vstore(out, random_in_unit_disk(state)); }

/* ===================================================================== */
/* Camera                                                                  */
/* ===================================================================== */

/* zmath normalize3: v / sqrt(dot3(v, v)), dot3 = (x*x + y*y) + z*z */
static v3 zm_normalize3(v3 v)
{
    float len = sqrtf((v.x * v.x + v.y * v.y) + v.z * v.z);
    return mk(v.x / len, v.y / len, v.z / len);
}

/* camera.zig:24-68 (host side, f32 via zmath) */
void orc_camera_init(const float lookfrom[3], const float lookat[3], const float vup[3],
                     float aspect_ratio, float vfov, float aperture, float focus_dist,
                     orc_camera *out)
{
    /* std.math.degreesToRadians(f32, a) = a * pi / 180 */
    float theta = vfov * PI_F / 180.0f;
    float h = (float)tan((double)(theta / 2.0f));
    float viewport_height = 2.0f * h;
    float viewport_width = aspect_ratio * viewport_height;
    v3 from = vload(lookfrom), at = vload(lookat), up = vload(vup);
    v3 w = zm_normalize3(vsub(from, at));
    v3 u = zm_normalize3(vcross(up, w));
    v3 v = vcross(w, u);
    v3 origin = from;
    v3 horizontal = vscale(focus_dist * viewport_width, u);
    v3 vertical = vscale(focus_dist * viewport_height, v);
    v3 llc = vsub(vsub(vsub(origin, mk(horizontal.x / 2.0f, horizontal.y / 2.0f, horizontal.z / 2.0f)),
                       mk(vertical.x / 2.0f, vertical.y / 2.0f, vertical.z / 2.0f)),
                  vscale(focus_dist, w));
    memset(out, 0, sizeof(*out));
    vstore(out->origin, origin);
    out->lens_radius = aperture / 2.0f;
    vstore(out->lower_left_corner, llc);
    vstore(out->horizontal, horizontal);
    vstore(out->vertical, vertical);
    vstore(out->u, u);
    vstore(out->v, v);
    vstore(out->w, w);
}

/* camera.wgsl:12-23: two disk draws even when lens_radius == 0; the
 * direction is not normalised. */
static void camera_get_ray(const orc_camera *cam, float s, float t, uint32_t *rng, v3 *o, v3 *d)
{
    v3 disk = random_in_unit_disk(rng);
    v3 rd = vscale(cam->lens_radius, disk);
    v3 cu = vload(cam->u), cv = vload(cam->v);
    v3 offset = vadd(vscale(rd.x, cu), vscale(rd.y, cv));
    v3 origin = vload(cam->origin);
    *o = vadd(origin, offset);
    v3 dir = vadd(vadd(vload(cam->lower_left_corner), vscale(s, vload(cam->horizontal))),
                  vscale(t, vload(cam->vertical)));
    *d = vsub(vsub(dir, origin), offset);
}

void orc_camera_get_ray(const orc_camera *cam, float s, float t, uint32_t *rng,
                        float origin[3], float direction[3])
{
    v3 o, d;
    camera_get_ray(cam, s, t, rng, &o, &d);
    vstore(origin, o);
    vstore(direction, d);
}

/* ===================================================================== */
/* Intersection: bvh.wgsl                                                  */
/* ===================================================================== */

/* bvh.wgsl:139-174 */
static inline float mycopysign(float a, float b) { return (b < 0.0f) ? -a : a; }

static v3 safe_invdir(v3 d)
{
    const float ooeps = 1e-5f;
    float x = (fabsf(d.x) > ooeps) ? d.x : mycopysign(ooeps, d.x);
    float y = (fabsf(d.y) > ooeps) ? d.y : mycopysign(ooeps, d.y);
    float z = (fabsf(d.z) > ooeps) ? d.z : mycopysign(ooeps, d.z);
    return mk(1.0f / x, 1.0f / y, 1.0f / z);
}

void orc_safe_invdir(const float d[3], float out[3]) { vstore(out, safe_invdir(vload(d))); }

/* bvh.wgsl:184-192 */
static inline int aabb_hit(const float *mn, const float *mx, v3 inv, v3 ox, float t_min,
                           float t_max, float *out_min, float *out_max)
{
    float fx = fmaf(mx[0], inv.x, ox.x), fy = fmaf(mx[1], inv.y, ox.y), fz = fmaf(mx[2], inv.z, ox.z);
    float nx = fmaf(mn[0], inv.x, ox.x), ny = fmaf(mn[1], inv.y, ox.y), nz = fmaf(mn[2], inv.z, ox.z);
    float tmax_x = omaxf(fx, nx), tmax_y = omaxf(fy, ny), tmax_z = omaxf(fz, nz);
    float tmin_x = ominf(fx, nx), tmin_y = ominf(fy, ny), tmin_z = ominf(fz, nz);
    float max_t = ominf(ominf(ominf(tmax_x, tmax_y), tmax_z), t_max);
    float min_t = omaxf(omaxf(omaxf(tmin_x, tmin_y), tmin_z), t_min);
    if (out_min) *out_min = min_t;
    if (out_max) *out_max = max_t;
    return min_t <= max_t;
}

void orc_aabb_hit(const float mn[3], const float mx[3], const float invdir[3],
                  const float oxinvdir[3], float t_min, float t_max, float out[2])
{
    aabb_hit(mn, mx, vload(invdir), vload(oxinvdir), t_min, t_max, &out[0], &out[1]);
}

/* bvh.wgsl:194-226: Moller-Trumbore, no determinant check, edges inclusive */
static float triangle_hit(v3 ro, v3 rd, v3 v1, v3 v2, v3 v3_, float t_min, float t_max, float *uv)
{
    v3 e1 = vsub(v2, v1);
    v3 e2 = vsub(v3_, v1);
    v3 s1 = vcross(rd, e2);
    float determinant = vdot(s1, e1);
    float invd = 1.0f / determinant;
    v3 d = vsub(ro, v1);
    float u = vdot(d, s1) * invd;
    if (u < 0.0f || u > 1.0f) return t_max;
    v3 s2 = vcross(d, e1);
    float v = vdot(rd, s2) * invd;
    if (v < 0.0f || u + v > 1.0f) return t_max;
    float t = vdot(e2, s2) * invd;
    if (t < t_min || t > t_max) return t_max;
    uv[0] = u;
    uv[1] = v;
    return t;
}

float orc_triangle_hit(const float o[3], const float d[3], const float v0[3], const float v1[3],
                       const float v2[3], float t_min, float t_max, float uv[2])
{
    return triangle_hit(vload(o), vload(d), vload(v0), vload(v1), vload(v2), t_min, t_max, uv);
}

/* bvh.wgsl:228-249: unit sphere at the origin */
static float sphere_hit(v3 ro, v3 rd, float t_min, float t_max)
{
    v3 oc = ro; /* origin - (0,0,0) */
    float a = rd.x * rd.x + rd.y * rd.y + rd.z * rd.z;
    float half_b = vdot(oc, rd);
    float c = (oc.x * oc.x + oc.y * oc.y + oc.z * oc.z) - 1.0f * 1.0f;
    float discriminant = half_b * half_b - a * c;
    if (discriminant < 0.0f) return t_max;
    float sqrtd = sqrtf(discriminant);
    float t = (-half_b - sqrtd) / a;
    if (t < t_min || t_max < t) {
        t = (-half_b + sqrtd) / a;
        if (t < t_min || t_max < t) return t_max;
    }
    return t;
}

float orc_sphere_hit(const float o[3], const float d[3], float t_min, float t_max)
{
    return sphere_hit(vload(o), vload(d), t_min, t_max);
}

/* transform.wgsl: the 16 floats are read as a column-major mat4x4, so column
 * c is floats [4c, 4c+4) and M * v = sum_c col_c * v_c. */
static inline v3 xform(const float *m, v3 p, float w)
{
    return mk(m[0] * p.x + m[4] * p.y + m[8] * p.z + m[12] * w,
              m[1] * p.x + m[5] * p.y + m[9] * p.z + m[13] * w,
              m[2] * p.x + m[6] * p.y + m[10] * p.z + m[14] * w);
}

/* transform.wgsl:16-19: transpose(M) * (n, 0) */
static inline v3 xform_normal(const float *m, v3 n)
{
    return mk(m[0] * n.x + m[1] * n.y + m[2] * n.z + m[3] * 0.0f,
              m[4] * n.x + m[5] * n.y + m[6] * n.z + m[7] * 0.0f,
              m[8] * n.x + m[9] * n.y + m[10] * n.z + m[11] * 0.0f);
}

#define FINISHED_TRAVERSE_BLAS 0xffffffffu
#define MAX_BVH_DEPTH 64

/* bvh.wgsl:14-137 */
static int bvh_hit(const orc_scene *sc, float t_min, v3 ro0, v3 rd0, orc_hit *out, orc_stats *st)
{
    float t_max = 3.40282e+38f;
    uint32_t node_stack[MAX_BVH_DEPTH];
    int stack_top = 0;
    uint32_t addr = sc->tlas_len - 1u;
    node_stack[stack_top] = addr;
    int traverse_tlas = 1;
    int hit_anything = 0;
    v3 ro = ro0, rd = rd0;
    v3 invdir = safe_invdir(rd);
    v3 oxinvdir = vmul(vneg(ro), invdir);
    const v3 nt_invdir = invdir;
    const v3 nt_oxinvdir = oxinvdir;
    uint32_t material_index = 0;
    uint32_t inverted_transform_id = 0;
    if (st) st->rays++;
    out->hit = 0;
    while (stack_top >= 0) {
        const orc_bvh_node *node = traverse_tlas ? &sc->tlas[addr] : &sc->blas[addr];
        if (st) st->node_visits++;
        switch (node->node_type) {
        case 0u: {
            int l = aabb_hit(node->lmin_or_v0, node->lmax_or_v1, invdir, oxinvdir, t_min, t_max, 0, 0);
            int r = aabb_hit(node->rmin_or_v2, node->rmax_or_v3, invdir, oxinvdir, t_min, t_max, 0, 0);
            if (l) {
                stack_top++;
                node_stack[stack_top] = node->left_or_custom_id;
            }
            if (r) {
                stack_top++;
                node_stack[stack_top] = node->right_or_material_index;
            }
            break;
        }
        case 1u: {
            inverted_transform_id = node->transform_id * 2u;
            const float *m = &sc->transforms[16u * inverted_transform_id];
            v3 to = xform(m, ro, 1.0f);
            v3 td = xform(m, rd, 0.0f);
            if (st) st->transform_fetches++;
            float t = sphere_hit(to, td, t_min, t_max);
            if (t < t_max) {
                hit_anything = 1;
                t_max = t;
                out->t = t;
                out->material_index = node->right_or_material_index;
                out->node_type = 1u;
                out->inverted_transform_id = inverted_transform_id;
            }
            break;
        }
        case 2u: {
            traverse_tlas = 0;
            stack_top++;
            node_stack[stack_top] = FINISHED_TRAVERSE_BLAS;
            stack_top++;
            node_stack[stack_top] = node->left_or_custom_id;
            inverted_transform_id = node->transform_id * 2u;
            material_index = node->right_or_material_index;
            const float *m = &sc->transforms[16u * inverted_transform_id];
            v3 to = xform(m, ro, 1.0f);
            v3 td = xform(m, rd, 0.0f);
            if (st) st->transform_fetches++;
            ro = to;
            rd = td;
            invdir = safe_invdir(rd);
            oxinvdir = vmul(vneg(ro), invdir);
            break;
        }
        case 3u: {
            float uv[2] = {0.0f, 0.0f};
            float t = triangle_hit(ro, rd, vload(node->lmin_or_v0), vload(node->lmax_or_v1),
                                   vload(node->rmin_or_v2), t_min, t_max, uv);
            if (t < t_max) {
                hit_anything = 1;
                t_max = t;
                out->t = t;
                out->material_index = material_index;
                out->node_type = 2u;
                out->inverted_transform_id = inverted_transform_id;
                out->tri_id = node->left_or_custom_id * 3u;
                out->uv[0] = uv[0];
                out->uv[1] = uv[1];
            }
            break;
        }
        default:
            break;
        }
        addr = node_stack[stack_top];
        stack_top--;
        if (addr == FINISHED_TRAVERSE_BLAS) {
            traverse_tlas = 1;
            ro = ro0;
            rd = rd0;
            invdir = nt_invdir;
            oxinvdir = nt_oxinvdir;
            addr = node_stack[stack_top];
            stack_top--;
        }
    }
    out->hit = (uint32_t)hit_anything;
    return hit_anything;
}

int orc_bvh_hit(const orc_scene *sc, float t_min, const float origin[3], const float direction[3],
                orc_hit *out, orc_stats *stats)
{
    memset(out, 0, sizeof(*out));
    return bvh_hit(sc, t_min, vload(origin), vload(direction), out, stats);
}

void orc_trace_rays(const orc_scene *sc, float t_min, const float *rays, uint32_t n, orc_hit *out,
                    int nthreads)
{
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
#endif
    for (int64_t i = 0; i < (int64_t)n; i++) {
        memset(&out[i], 0, sizeof(orc_hit));
        bvh_hit(sc, t_min, vload(&rays[6 * i]), vload(&rays[6 * i + 3]), &out[i], 0);
    }
    (void)nthreads;
}

/* ===================================================================== */
/* FP64 brute-force closest hit                                            */
/* ===================================================================== */

typedef struct {
    double best_t, second_t;
    uint32_t best_tid, best_tri;
    int ambiguous;
    double uncertain_t; /* smallest t of a primitive that may or may not be hit */
    uint32_t best_leaf, best_node; /* TLAS leaf and BLAS triangle node of the best hit */
} bf_state;

static void bf_consider(bf_state *s, double t, uint32_t tid, uint32_t tri, int certain,
                        uint32_t leaf, uint32_t node)
{
    if (!certain) {
        if (t < s->uncertain_t) s->uncertain_t = t;
        return;
    }
    if (t < s->best_t) {
        s->second_t = s->best_t;
        s->best_t = t;
        s->best_tid = tid;
        s->best_tri = tri;
        s->best_leaf = leaf;
        s->best_node = node;
    } else if (t < s->second_t) {
        s->second_t = t;
    }
}

/* First-order forward error bound, relative to t, of the reference's FP32
 * operation sequence for the best hit (transform.wgsl:7-14 then bvh.wgsl
 * sphere_hit or triangle_hit), evaluated in FP64 with u = 2^-24 per
 * rounding. It measures how well FP32 can resolve t for this ray at all:
 * grazing hits, origins close to a huge sphere's surface and cancellation
 * in the Moller-Trumbore dot products all show up here. */
#define BF_U 5.9604644775390625e-08

static void bf_xform_err(const float *m, const double *p, double w, double *val, double *err)
{
    for (int i = 0; i < 3; i++) {
        double t0 = (double)m[i] * p[0], t1 = (double)m[4 + i] * p[1], t2 = (double)m[8 + i] * p[2],
               t3 = (double)m[12 + i] * w;
        val[i] = t0 + t1 + t2 + t3;
        err[i] = 4.0 * BF_U * (fabs(t0) + fabs(t1) + fabs(t2) + fabs(t3));
    }
}

static void bf_cross_err(const double *a, const double *ea, const double *b, const double *eb,
                         double *c, double *ec)
{
    static const int I[3] = {1, 2, 0}, J[3] = {2, 0, 1};
    for (int k = 0; k < 3; k++) {
        int i = I[k], j = J[k];
        c[k] = a[i] * b[j] - a[j] * b[i];
        ec[k] = 2.0 * BF_U * (fabs(a[i] * b[j]) + fabs(a[j] * b[i])) + fabs(b[j]) * ea[i] +
                fabs(a[i]) * eb[j] + fabs(b[i]) * ea[j] + fabs(a[j]) * eb[i];
    }
}

static double bf_dot_err(const double *a, const double *ea, const double *b, const double *eb,
                         double *val)
{
    double v = 0, e = 0;
    for (int k = 0; k < 3; k++) {
        v += a[k] * b[k];
        e += 3.0 * BF_U * fabs(a[k] * b[k]) + fabs(a[k]) * eb[k] + fabs(b[k]) * ea[k];
    }
    *val = v;
    return e;
}

static double bf_fp32_bound(const orc_scene *sc, const bf_state *s, const double *o, const double *d,
                            double t_min)
{
    const orc_bvh_node *nd = &sc->tlas[s->best_leaf];
    const float *m = &sc->transforms[16u * 2u * nd->transform_id];
    double lo[3], elo[3], ld[3], eld[3];
    bf_xform_err(m, o, 1.0, lo, elo);
    bf_xform_err(m, d, 0.0, ld, eld);
    if (nd->node_type == 1u) {
        double a, hb, oo;
        double ea = bf_dot_err(ld, eld, ld, eld, &a);
        double ehb = bf_dot_err(lo, elo, ld, eld, &hb);
        double eoo = bf_dot_err(lo, elo, lo, elo, &oo);
        double c = oo - 1.0, ec = eoo + BF_U * (fabs(oo) + 1.0);
        double disc = hb * hb - a * c;
        double edisc = 2.0 * BF_U * (hb * hb + fabs(a * c)) + 2.0 * fabs(hb) * ehb + fabs(c) * ea +
                       fabs(a) * ec;
        if (disc <= 0.0) return INFINITY;
        double sq = sqrt(disc), esq = BF_U * sq + edisc / (2.0 * sq);
        double t = (-hb - sq) / a;
        if (t < t_min) t = (-hb + sq) / a;
        double et = (ehb + esq) / fabs(a) + fabs(t) * ea / fabs(a) + 2.0 * BF_U * fabs(t);
        return et / fabs(t);
    }
    const orc_bvh_node *tn = &sc->blas[s->best_node];
    double v0[3], e1[3], e2[3], ee1[3], ee2[3], dd[3], edd[3];
    for (int j = 0; j < 3; j++) {
        v0[j] = tn->lmin_or_v0[j];
        e1[j] = (double)tn->lmax_or_v1[j] - v0[j];
        e2[j] = (double)tn->rmin_or_v2[j] - v0[j];
        ee1[j] = BF_U * fabs(e1[j]);
        ee2[j] = BF_U * fabs(e2[j]);
        dd[j] = lo[j] - v0[j];
        edd[j] = elo[j] + BF_U * fabs(dd[j]);
    }
    double s1[3], es1[3], s2[3], es2[3], det, num;
    bf_cross_err(ld, eld, e2, ee2, s1, es1);
    double edet = bf_dot_err(s1, es1, e1, ee1, &det);
    bf_cross_err(dd, edd, e1, ee1, s2, es2);
    double enum_ = bf_dot_err(e2, ee2, s2, es2, &num);
    if (det == 0.0 || num == 0.0) return INFINITY;
    return enum_ / fabs(num) + edet / fabs(det) + 3.0 * BF_U;
}

static void bf_xform(const float *m, const double *p, double w, double *out)
{
    for (int i = 0; i < 3; i++)
        out[i] = (double)m[i] * p[0] + (double)m[4 + i] * p[1] + (double)m[8 + i] * p[2] +
                 (double)m[12 + i] * w;
}

static void collect_triangles(const orc_scene *sc, uint32_t root, uint32_t **list, uint32_t *len,
                              uint32_t *cap)
{
    uint32_t stack[128];
    int top = 0;
    stack[0] = root;
    while (top >= 0) {
        uint32_t a = stack[top--];
        const orc_bvh_node *nd = &sc->blas[a];
        if (nd->node_type == 3u) {
            if (*len == *cap) {
                *cap = *cap ? *cap * 2 : 64;
                *list = (uint32_t *)realloc(*list, *cap * sizeof(uint32_t));
            }
            (*list)[(*len)++] = a;
        } else if (nd->node_type == 0u && top < 126) {
            stack[++top] = nd->left_or_custom_id;
            stack[++top] = nd->right_or_material_index;
        }
    }
}

void orc_brute_force(const orc_scene *sc, double t_min, const float *rays, uint32_t n,
                     double rel_eps, double *out8, int nthreads)
{
    /* leaves of the TLAS and, per mesh leaf, its triangle nodes */
    uint32_t nleaf = 0;
    uint32_t *leaf = (uint32_t *)malloc(sizeof(uint32_t) * (sc->tlas_len + 1));
    uint32_t **tris = (uint32_t **)calloc(sc->tlas_len + 1, sizeof(uint32_t *));
    uint32_t *ntris = (uint32_t *)calloc(sc->tlas_len + 1, sizeof(uint32_t));
    for (uint32_t i = 0; i < sc->tlas_len; i++) {
        const orc_bvh_node *nd = &sc->tlas[i];
        if (nd->node_type == 1u || nd->node_type == 2u) {
            if (nd->node_type == 2u) {
                uint32_t cap = 0;
                collect_triangles(sc, nd->left_or_custom_id, &tris[nleaf], &ntris[nleaf], &cap);
            }
            leaf[nleaf++] = i;
        }
    }
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
#endif
    for (int64_t r = 0; r < (int64_t)n; r++) {
        const double eps = rel_eps;
        double o[3] = {rays[6 * r], rays[6 * r + 1], rays[6 * r + 2]};
        double d[3] = {rays[6 * r + 3], rays[6 * r + 4], rays[6 * r + 5]};
        bf_state s;
        s.best_t = INFINITY;
        s.second_t = INFINITY;
        s.best_tid = 0xffffffffu;
        s.best_tri = 0xffffffffu;
        s.ambiguous = 0;
        s.uncertain_t = INFINITY;
        s.best_leaf = s.best_node = 0;
        for (uint32_t k = 0; k < nleaf; k++) {
            const orc_bvh_node *nd = &sc->tlas[leaf[k]];
            const float *m = &sc->transforms[16u * 2u * nd->transform_id];
            double lo[3], ld[3];
            bf_xform(m, o, 1.0, lo);
            bf_xform(m, d, 0.0, ld);
            if (nd->node_type == 1u) {
                double a = ld[0] * ld[0] + ld[1] * ld[1] + ld[2] * ld[2];
                double hb = lo[0] * ld[0] + lo[1] * ld[1] + lo[2] * ld[2];
                double c = lo[0] * lo[0] + lo[1] * lo[1] + lo[2] * lo[2] - 1.0;
                double disc = hb * hb - a * c;
                double scale = hb * hb + fabs(a * c) + 1e-300;
                if (disc < -eps * scale) continue;
                int graze = fabs(disc) <= eps * scale;
                double sq = sqrt(disc > 0 ? disc : 0);
                double roots[2] = {(-hb - sq) / a, (-hb + sq) / a};
                for (int q = 0; q < 2; q++) {
                    double t = roots[q];
                    if (t < t_min * (1.0 - eps)) continue;
                    int certain = !graze && t > t_min * (1.0 + eps);
                    bf_consider(&s, t, nd->transform_id, 0xffffffffu, certain, leaf[k], 0);
                    break; /* nearest admissible root only */
                }
            } else {
                for (uint32_t q = 0; q < ntris[k]; q++) {
                    const orc_bvh_node *tn = &sc->blas[tris[k][q]];
                    double v0[3], e1[3], e2[3];
                    for (int j = 0; j < 3; j++) {
                        v0[j] = tn->lmin_or_v0[j];
                        e1[j] = (double)tn->lmax_or_v1[j] - v0[j];
                        e2[j] = (double)tn->rmin_or_v2[j] - v0[j];
                    }
                    double s1[3] = {ld[1] * e2[2] - ld[2] * e2[1], ld[2] * e2[0] - ld[0] * e2[2],
                                    ld[0] * e2[1] - ld[1] * e2[0]};
                    double det = s1[0] * e1[0] + s1[1] * e1[1] + s1[2] * e1[2];
                    double dn = sqrt(ld[0] * ld[0] + ld[1] * ld[1] + ld[2] * ld[2]);
                    double en = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]) *
                                sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
                    if (fabs(det) <= 1e-9 * dn * en) continue; /* parallel: treated as miss */
                    double dd[3] = {lo[0] - v0[0], lo[1] - v0[1], lo[2] - v0[2]};
                    double u = (dd[0] * s1[0] + dd[1] * s1[1] + dd[2] * s1[2]) / det;
                    double s2[3] = {dd[1] * e1[2] - dd[2] * e1[1], dd[2] * e1[0] - dd[0] * e1[2],
                                    dd[0] * e1[1] - dd[1] * e1[0]};
                    double v = (ld[0] * s2[0] + ld[1] * s2[1] + ld[2] * s2[2]) / det;
                    double t = (e2[0] * s2[0] + e2[1] * s2[1] + e2[2] * s2[2]) / det;
                    double margin = u;
                    if (v < margin) margin = v;
                    if (1.0 - u - v < margin) margin = 1.0 - u - v;
                    if (margin < -eps) continue;
                    if (t < t_min * (1.0 - eps)) continue;
                    int certain = margin > eps && t > t_min * (1.0 + eps);
                    bf_consider(&s, t, nd->transform_id, tn->left_or_custom_id, certain, leaf[k],
                                tris[k][q]);
                }
            }
        }
        int amb = 0;
        if (!isinf(s.uncertain_t) && s.uncertain_t <= s.best_t * (1.0 + eps)) amb = 1;
        if (!isinf(s.best_t) && s.second_t <= s.best_t * (1.0 + eps)) amb = 1;
        double *w = &out8[8 * r];
        w[0] = isinf(s.best_t) ? 0.0 : 1.0;
        w[1] = s.best_t;
        w[2] = (double)s.best_tid;
        w[3] = (double)s.best_tri;
        w[4] = (double)amb;
        w[5] = s.second_t;
        w[6] = 0.0;
        w[7] = 0.0;
        if (!isinf(s.best_t)) {
            w[6] = bf_fp32_bound(sc, &s, o, d, t_min);
        }
    }
    for (uint32_t k = 0; k < nleaf; k++) free(tris[k]);
    free(tris);
    free(ntris);
    free(leaf);
    (void)nthreads;
}

/* ===================================================================== */
/* Shading: material.wgsl, hitrecord.wgsl, pathtracer.wgsl                 */
/* ===================================================================== */

typedef struct {
    v3 p, normal;
    float uv[2];
    uint32_t material_index;
    float t;
    int front_face;
} hit_record;

/* hitrecord.wgsl:10-18 */
static void set_face_normal(hit_record *h, v3 rd, v3 outward)
{
    if (vdot(rd, outward) > 0.0f) {
        h->normal = vneg(outward);
        h->front_face = 0;
    } else {
        h->normal = outward;
        h->front_face = 1;
    }
}

static inline float half_to_float(uint16_t h)
{
    uint32_t sign = (uint32_t)(h >> 15) << 31;
    uint32_t exp = (h >> 10) & 0x1fu;
    uint32_t man = h & 0x3ffu;
    if (exp == 0) {
        float f = (float)man * (1.0f / 16777216.0f); /* man * 2^-24 */
        return sign ? -f : f;
    }
    if (exp == 31) return u2f(sign | 0x7f800000u | (man << 13));
    return u2f(sign | ((exp + 112u) << 23) | (man << 13));
}

/* textureSampleLevel with the default sampler (webgpu.zig:812-825):
 * nearest filtering, clamp-to-edge, level 0; rgba8unorm is byte / 255. */
static v3 sample_texture(const orc_texture *tx, const float *uv)
{
    float fx = floorf(uv[0] * (float)tx->width);
    float fy = floorf(uv[1] * (float)tx->height);
    fx = ominf(omaxf(fx, 0.0f), (float)(tx->width - 1u));
    fy = ominf(omaxf(fy, 0.0f), (float)(tx->height - 1u));
    uint32_t ix = (uint32_t)fx, iy = (uint32_t)fy;
    uint32_t nc = tx->num_components;
    size_t texel = (size_t)iy * tx->width + ix;
    float c[4] = {0.0f, 0.0f, 0.0f, 1.0f};
    if (tx->is_hdr) {
        const uint16_t *d = (const uint16_t *)tx->data + texel * nc;
        for (uint32_t k = 0; k < nc && k < 3; k++) c[k] = half_to_float(d[k]);
    } else {
        const uint8_t *d = tx->data + texel * nc;
        for (uint32_t k = 0; k < nc && k < 3; k++) c[k] = (float)d[k] / 255.0f;
    }
    return mk(c[0], c[1], c[2]);
}

/* material.wgsl:91-99 */
static v3 material_get_color(const orc_scene *sc, const orc_constant_params *p, const orc_material *m,
                             const float *uv, orc_stats *st)
{
    if (m->albedo_texture_index < p->textures_count) {
        if (st) st->textured_hits++;
        return sample_texture(&sc->textures[m->albedo_texture_index], uv);
    }
    return vload(m->albedo);
}

static inline int near_zero(v3 e)
{
    const float s = 1e-8f;
    return fabsf(e.x) < s && fabsf(e.y) < s && fabsf(e.z) < s;
}

/* WGSL reflect(e1, e2) = e1 - 2 * dot(e2, e1) * e2 */
static inline v3 reflect_(v3 v, v3 n) { return vsub(v, vscale(2.0f * vdot(n, v), n)); }

/* WGSL refract(e1, e2, e3) */
static inline v3 refract_(v3 e1, v3 e2, float eta)
{
    float d = vdot(e2, e1);
    float k = 1.0f - eta * eta * (1.0f - d * d);
    if (k < 0.0f) return mk(0.0f, 0.0f, 0.0f);
    return vsub(vscale(eta, e1), vscale(eta * d + sqrtf(k), e2));
}

/* material.wgsl:29-34 */
static inline float reflectance(float cosine, float ref_idx)
{
    float r0 = (1.0f - ref_idx) / (1.0f + ref_idx);
    r0 = r0 * r0;
    return r0 + (1.0f - r0) * orc_powf(1.0f - cosine, 5.0f);
}

/* material.wgsl:61-77; returns 1 when the path continues */
static int material_scatter(const orc_scene *sc, const orc_constant_params *p, v3 rd,
                            const hit_record *h, uint32_t *rng, v3 *att, v3 *so, v3 *sd,
                            orc_stats *st)
{
    const orc_material *m = &sc->materials[h->material_index];
    switch (m->type) {
    case 0u: { /* material.wgsl:9-20 */
        v3 dir = vadd(h->normal, random_unit_vector(rng));
        if (near_zero(dir)) dir = h->normal;
        *so = h->p;
        *sd = dir;
        *att = material_get_color(sc, p, m, h->uv, st);
        return 1;
    }
    case 1u: { /* material.wgsl:22-27 */
        v3 dir = vadd(reflect_(vnormalize(rd), h->normal), vscale(m->fuzz, random_in_unit_sphere(rng)));
        *so = h->p;
        *sd = dir;
        *att = material_get_color(sc, p, m, h->uv, st);
        return vdot(dir, h->normal) > 0.0f;
    }
    case 2u: { /* material.wgsl:36-59 */
        *att = mk(1.0f, 1.0f, 1.0f);
        float ratio = h->front_face ? (1.0f / m->ior) : m->ior;
        v3 ud = vnormalize(rd);
        float cos_theta = ominf(vdot(vneg(ud), h->normal), 1.0f);
        float sin_theta = sqrtf(1.0f - cos_theta * cos_theta);
        int cannot_refract = ratio * sin_theta > 1.0f;
        v3 dir;
        /* `||` short-circuits: the draw happens only when refraction is possible */
        if (cannot_refract || reflectance(cos_theta, ratio) > orc_random_f32(rng))
            dir = reflect_(ud, h->normal);
        else
            dir = refract_(ud, h->normal, ratio);
        *so = h->p;
        *sd = dir;
        return 1;
    }
    default:
        return 0;
    }
}

/* material.wgsl:79-89 */
static v3 material_emit(const orc_scene *sc, const orc_constant_params *p, const hit_record *h,
                        orc_stats *st)
{
    const orc_material *m = &sc->materials[h->material_index];
    if (m->type == 3u) return material_get_color(sc, p, m, h->uv, st);
    return mk(0.0f, 0.0f, 0.0f);
}

/* pathtracer.wgsl:111-144: the hit record of a closest hit */
static void make_hit_record(const orc_scene *sc, v3 ro, v3 rd, const orc_hit *hh, hit_record *hit,
                            orc_stats *st)
{
    uint32_t transform_id = hh->inverted_transform_id + 1u;
    memset(hit, 0, sizeof(*hit));
    hit->t = hh->t;
    hit->p = vadd(ro, vscale(hh->t, rd));
    hit->material_index = hh->material_index;
    if (hh->node_type == 1u) {
        if (st) { st->sphere_hits++; }
        v3 center = xform(&sc->transforms[16u * transform_id], mk(0.0f, 0.0f, 0.0f), 1.0f);
        v3 on = vnormalize(vsub(hit->p, center));
        float theta = orc_acosf(-on.y);
        float phi = orc_atan2f(-on.z, on.x) + PI_F;
        hit->uv[0] = phi / (2.0f * PI_F);
        hit->uv[1] = theta / PI_F;
        set_face_normal(hit, rd, on);
    } else if (hh->node_type == 2u) {
        if (st) { st->mesh_hits++; }
        uint32_t tri = hh->tri_id;
        v3 n0 = vload(&sc->normals[4u * sc->normal_indices[tri]]);
        v3 n1 = vload(&sc->normals[4u * sc->normal_indices[tri + 1u]]);
        v3 n2 = vload(&sc->normals[4u * sc->normal_indices[tri + 2u]]);
        const float *uv0 = &sc->uvs[2u * sc->uv_indices[tri]];
        const float *uv1 = &sc->uvs[2u * sc->uv_indices[tri + 1u]];
        const float *uv2 = &sc->uvs[2u * sc->uv_indices[tri + 2u]];
        float w = 1.0f - hh->uv[0] - hh->uv[1];
        v3 normal = vadd(vadd(vscale(w, n0), vscale(hh->uv[0], n1)), vscale(hh->uv[1], n2));
        hit->uv[0] = w * uv0[0] + hh->uv[0] * uv1[0] + hh->uv[1] * uv2[0];
        hit->uv[1] = w * uv0[1] + hh->uv[0] * uv1[1] + hh->uv[1] * uv2[1];
        v3 on = vnormalize(xform_normal(&sc->transforms[16u * hh->inverted_transform_id], normal));
        set_face_normal(hit, rd, on);
    }
}

/* pathtracer.wgsl:95-161: depth exhaustion returns the throughput */
static v3 ray_color(const orc_scene *sc, const orc_constant_params *p, v3 ro, v3 rd, uint32_t *rng,
                    orc_stats *st)
{
    v3 final_color = mk(1.0f, 1.0f, 1.0f);
    for (uint32_t i = 0; i < p->depth; i++) {
        orc_hit hh;
        memset(&hh, 0, sizeof(hh));
        if (!bvh_hit(sc, p->ray_cast_epsilon, ro, rd, &hh, st)) {
            v3 ud = vnormalize(rd);
            float tt = 0.5f * (ud.y + 1.0f);
            float a = (1.0f - tt) * 1.0f;
            final_color = vmul(final_color, mk(a + tt * 0.5f, a + tt * 0.7f, a + tt * 1.0f));
            break;
        }
        hit_record hit;
        make_hit_record(sc, ro, rd, &hh, &hit, st);
        v3 att, so, sd;
        if (material_scatter(sc, p, rd, &hit, rng, &att, &so, &sd, st)) {
            ro = so;
            rd = sd;
            final_color = vmul(final_color, att);
        } else {
            final_color = vmul(final_color, material_emit(sc, p, &hit, st));
            break;
        }
    }
    return final_color;
}

/* ---- exported pieces of ray_color (closed-form tests) ---- */
static void record_out(const hit_record *h, orc_hit_record *o)
{
    vstore(o->p, h->p);
    vstore(o->normal, h->normal);
    o->uv[0] = h->uv[0];
    o->uv[1] = h->uv[1];
    o->material_index = h->material_index;
    o->t = h->t;
    o->front_face = h->front_face;
}

static void record_in(const orc_hit_record *o, hit_record *h)
{
    memset(h, 0, sizeof(*h));
    h->p = vload(o->p);
    h->normal = vload(o->normal);
    h->uv[0] = o->uv[0];
    h->uv[1] = o->uv[1];
    h->material_index = o->material_index;
    h->t = o->t;
    h->front_face = o->front_face;
}

void orc_make_hit_record(const orc_scene *sc, const float ro[3], const float rd[3], const orc_hit *hit,
                         orc_hit_record *out)
{
    hit_record h;
    make_hit_record(sc, vload(ro), vload(rd), hit, &h, NULL);
    record_out(&h, out);
}

int orc_material_scatter(const orc_scene *sc, const orc_constant_params *p, const float rd[3],
                         const orc_hit_record *h, uint32_t *rng, float att[3], float origin[3],
                         float direction[3])
{
    hit_record hr;
    record_in(h, &hr);
    v3 a = mk(0.0f, 0.0f, 0.0f), so = a, sd = a;
    int r = material_scatter(sc, p, vload(rd), &hr, rng, &a, &so, &sd, NULL);
    vstore(att, a);
    vstore(origin, so);
    vstore(direction, sd);
    return r;
}

void orc_material_emit(const orc_scene *sc, const orc_constant_params *p, const orc_hit_record *h,
                       float out[3])
{
    hit_record hr;
    record_in(h, &hr);
    vstore(out, material_emit(sc, p, &hr, NULL));
}

void orc_sample_texture(const orc_texture *tx, const float uv[2], float out[3])
{
    vstore(out, sample_texture(tx, uv));
}

float orc_half_to_float(uint16_t h) { return half_to_float(h); }

/* pathtracer.wgsl:64-79 (one sample) */
static v3 sample_pixel(const orc_scene *sc, const orc_constant_params *p, uint32_t x, uint32_t y,
                       uint32_t *rng, orc_stats *st)
{
    float u = ((float)x + orc_random_f32(rng)) / (float)(p->width - 1u);
    float v = ((float)y + orc_random_f32(rng)) / (float)(p->height - 1u);
    v3 o, d;
    camera_get_ray(&p->camera, u, v, rng, &o, &d);
    if (st) st->samples++;
    return ray_color(sc, p, o, d, rng, st);
}

/* pathtracer.wgsl:81-93, returns the framebuffer row index offset */
static void post_process(const orc_constant_params *p, const float *acc, float cur, float *rgba)
{
    for (int k = 0; k < 4; k++) {
        float c = acc[k] / cur;
        c = ominf(omaxf(c, 0.0f), 1.0f);
        rgba[k] = c;
    }
    for (int k = 0; k < 3; k++) rgba[k] = orc_powf(rgba[k], p->inverted_gamma);
}

/* One pixel through `iterations` updates (path_tracer.zig:212-265). */
static void render_one(const orc_scene *sc, const orc_constant_params *p, uint32_t iterations,
                       uint32_t pixel, float *acc, uint32_t *rng, float *rgba_out, orc_stats *st)
{
    uint32_t x = pixel % p->width, y = pixel / p->width;
    float cur = p->current_iteration;
    uint32_t n = iterations > 1u ? iterations : 1u; /* iterations <= 1 runs the fused `main` once */
    for (uint32_t it = 0; it < n; it++) {
        cur = cur + 1.0f; /* State.nextIteration (state.zig:69) */
        v3 c = sample_pixel(sc, p, x, y, rng, st);
        if (cur > 1.0f) {
            acc[0] = acc[0] + c.x;
            acc[1] = acc[1] + c.y;
            acc[2] = acc[2] + c.z;
            acc[3] = acc[3] + 1.0f;
        } else {
            acc[0] = c.x;
            acc[1] = c.y;
            acc[2] = c.z;
            acc[3] = 1.0f;
        }
    }
    post_process(p, acc, cur, rgba_out);
}

static float advance(float cur, uint32_t iterations)
{
    uint32_t n = iterations > 1u ? iterations : 1u;
    for (uint32_t i = 0; i < n; i++) cur = cur + 1.0f;
    return cur;
}

static void merge_stats(orc_stats *dst, const orc_stats *src)
{
    dst->rays += src->rays;
    dst->node_visits += src->node_visits;
    dst->transform_fetches += src->transform_fetches;
    dst->sphere_hits += src->sphere_hits;
    dst->mesh_hits += src->mesh_hits;
    dst->textured_hits += src->textured_hits;
    dst->samples += src->samples;
}

float orc_render(const orc_scene *sc, const orc_constant_params *params, uint32_t iterations,
                 float *acc, uint32_t *rng, float *fb, int nthreads, orc_stats *stats)
{
    const uint32_t n = params->width * params->height;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
    {
        orc_stats local;
        memset(&local, 0, sizeof(local));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
        for (int64_t i = 0; i < (int64_t)n; i++) {
            float rgba[4];
            render_one(sc, params, iterations, (uint32_t)i, &acc[4 * i], &rng[i], rgba,
                       stats ? &local : 0);
            uint32_t x = (uint32_t)i % params->width, y = (uint32_t)i / params->width;
            uint32_t dst = params->flip_y < 1u ? (uint32_t)i
                                               : params->width * (params->height - y - 1u) + x;
            memcpy(&fb[4 * (size_t)dst], rgba, sizeof(rgba));
        }
        if (stats) {
#ifdef _OPENMP
#pragma omp critical
#endif
            merge_stats(stats, &local);
        }
    }
    (void)nthreads;
    return advance(params->current_iteration, iterations);
}

float orc_render_pixels(const orc_scene *sc, const orc_constant_params *params, uint32_t iterations,
                        const uint32_t *pixels, uint32_t n, float *acc, uint32_t *rng, float *fb,
                        int nthreads, orc_stats *stats)
{
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
    {
        orc_stats local;
        memset(&local, 0, sizeof(local));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
        for (int64_t i = 0; i < (int64_t)n; i++)
            render_one(sc, params, iterations, pixels[i], &acc[4 * i], &rng[i], &fb[4 * i],
                       stats ? &local : 0);
        if (stats) {
#ifdef _OPENMP
#pragma omp critical
#endif
            merge_stats(stats, &local);
        }
    }
    (void)nthreads;
    return advance(params->current_iteration, iterations);
}

void orc_seed_rng(uint32_t *rng, uint32_t pixel_count, uint32_t seed)
{
    for (uint32_t i = 0; i < pixel_count; i++) rng[i] = i + seed * pixel_count;
}

/* ===================================================================== */
/* zmath (row-vector matrices, rows contiguous), f32 without FMA           */
/* ===================================================================== */

void orc_mat_identity(float m[16])
{
    memset(m, 0, 16 * sizeof(float));
    m[0] = m[5] = m[10] = m[15] = 1.0f;
}

/* zmath mulMat: row = (x*b0 + z*b2) + (y*b1 + w*b3) */
void orc_mat_mul(const float a[16], const float b[16], float out[16])
{
    float r[16];
    for (int i = 0; i < 4; i++) {
        float x = a[4 * i], y = a[4 * i + 1], z = a[4 * i + 2], w = a[4 * i + 3];
        for (int j = 0; j < 4; j++)
            r[4 * i + j] = (x * b[j] + z * b[8 + j]) + (y * b[4 + j] + w * b[12 + j]);
    }
    memcpy(out, r, sizeof(r));
}

void orc_mat_scaling(float sx, float sy, float sz, float out[16])
{
    orc_mat_identity(out);
    out[0] = sx;
    out[5] = sy;
    out[10] = sz;
}

void orc_mat_translation(float x, float y, float z, float out[16])
{
    orc_mat_identity(out);
    out[12] = x;
    out[13] = y;
    out[14] = z;
}

/* zmath matFromQuat (DirectXMath XMMatrixRotationQuaternion order) */
void orc_mat_from_quat(const float q[4], float out[16])
{
    float x = q[0], y = q[1], z = q[2], w = q[3];
    float x2 = x + x, y2 = y + y, z2 = z + z;
    float xx = x * x2, yy = y * y2, zz = z * z2;
    float r0x = (1.0f - yy) - zz, r0y = (1.0f - xx) - zz, r0z = (1.0f - xx) - yy;
    float v0x = x * z2, v0y = x * y2, v0z = y * z2;
    float v1x = w * y2, v1y = w * z2, v1z = w * x2;
    float r1x = v0x + v1x, r1y = v0y + v1y, r1z = v0z + v1z;
    float r2x = v0x - v1x, r2y = v0y - v1y, r2z = v0z - v1z;
    float m[16] = {r0x, r1y, r2x, 0.0f, r2y, r0y, r1z, 0.0f, r1x, r2z, r0z, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f};
    memcpy(out, m, sizeof(m));
}

/* Inverse in double through the cofactor expansion, rounded once to f32 and
 * written without negative zeros. The f32 result is the correctly rounded
 * inverse except in double-rounding cases; exact zeros stay exact. */
int orc_mat_inverse(const float m[16], float out[16])
{
    double a[16], inv[16];
    for (int i = 0; i < 16; i++) a[i] = m[i];
    inv[0] = a[5] * a[10] * a[15] - a[5] * a[11] * a[14] - a[9] * a[6] * a[15] + a[9] * a[7] * a[14] +
             a[13] * a[6] * a[11] - a[13] * a[7] * a[10];
    inv[4] = -a[4] * a[10] * a[15] + a[4] * a[11] * a[14] + a[8] * a[6] * a[15] - a[8] * a[7] * a[14] -
             a[12] * a[6] * a[11] + a[12] * a[7] * a[10];
    inv[8] = a[4] * a[9] * a[15] - a[4] * a[11] * a[13] - a[8] * a[5] * a[15] + a[8] * a[7] * a[13] +
             a[12] * a[5] * a[11] - a[12] * a[7] * a[9];
    inv[12] = -a[4] * a[9] * a[14] + a[4] * a[10] * a[13] + a[8] * a[5] * a[14] - a[8] * a[6] * a[13] -
              a[12] * a[5] * a[10] + a[12] * a[6] * a[9];
    inv[1] = -a[1] * a[10] * a[15] + a[1] * a[11] * a[14] + a[9] * a[2] * a[15] - a[9] * a[3] * a[14] -
             a[13] * a[2] * a[11] + a[13] * a[3] * a[10];
    inv[5] = a[0] * a[10] * a[15] - a[0] * a[11] * a[14] - a[8] * a[2] * a[15] + a[8] * a[3] * a[14] +
             a[12] * a[2] * a[11] - a[12] * a[3] * a[10];
    inv[9] = -a[0] * a[9] * a[15] + a[0] * a[11] * a[13] + a[8] * a[1] * a[15] - a[8] * a[3] * a[13] -
             a[12] * a[1] * a[11] + a[12] * a[3] * a[9];
    inv[13] = a[0] * a[9] * a[14] - a[0] * a[10] * a[13] - a[8] * a[1] * a[14] + a[8] * a[2] * a[13] +
              a[12] * a[1] * a[10] - a[12] * a[2] * a[9];
    inv[2] = a[1] * a[6] * a[15] - a[1] * a[7] * a[14] - a[5] * a[2] * a[15] + a[5] * a[3] * a[14] +
             a[13] * a[2] * a[7] - a[13] * a[3] * a[6];
    inv[6] = -a[0] * a[6] * a[15] + a[0] * a[7] * a[14] + a[4] * a[2] * a[15] - a[4] * a[3] * a[14] -
             a[12] * a[2] * a[7] + a[12] * a[3] * a[6];
    inv[10] = a[0] * a[5] * a[15] - a[0] * a[7] * a[13] - a[4] * a[1] * a[15] + a[4] * a[3] * a[13] +
              a[12] * a[1] * a[7] - a[12] * a[3] * a[5];
    inv[14] = -a[0] * a[5] * a[14] + a[0] * a[6] * a[13] + a[4] * a[1] * a[14] - a[4] * a[2] * a[13] -
              a[12] * a[1] * a[6] + a[12] * a[2] * a[5];
    inv[3] = -a[1] * a[6] * a[11] + a[1] * a[7] * a[10] + a[5] * a[2] * a[11] - a[5] * a[3] * a[10] -
             a[9] * a[2] * a[7] + a[9] * a[3] * a[6];
    inv[7] = a[0] * a[6] * a[11] - a[0] * a[7] * a[10] - a[4] * a[2] * a[11] + a[4] * a[3] * a[10] +
             a[8] * a[2] * a[7] - a[8] * a[3] * a[6];
    inv[11] = -a[0] * a[5] * a[11] + a[0] * a[7] * a[9] + a[4] * a[1] * a[11] - a[4] * a[3] * a[9] -
              a[8] * a[1] * a[7] + a[8] * a[3] * a[5];
    inv[15] = a[0] * a[5] * a[10] - a[0] * a[6] * a[9] - a[4] * a[1] * a[10] + a[4] * a[2] * a[9] +
              a[8] * a[1] * a[6] - a[8] * a[2] * a[5];
    double det = a[0] * inv[0] + a[1] * inv[4] + a[2] * inv[8] + a[3] * inv[12];
    if (det == 0.0) return -1;
    for (int i = 0; i < 16; i++) out[i] = (float)(inv[i] / det) + 0.0f; /* no negative zeros */
    return 0;
}

/* zmath vecMulMat: ((x*m0 + y*m1) + z*m2) + w*m3 */
void orc_vec_mul_mat(const float v[4], const float m[16], float out[4])
{
    float r[4];
    for (int j = 0; j < 4; j++)
        r[j] = ((v[0] * m[j] + v[1] * m[4 + j]) + v[2] * m[8 + j]) + v[3] * m[12 + j];
    memcpy(out, r, sizeof(r));
}

static float zm_dot3(const float *a, const float *b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

static int approx_eql(float a, float b) { return fabsf(a - b) <= 0.0000001f; }

/* math.zig:13-41 */
void orc_rotation_between_vectors(const float a[4], const float b[4], float out[16])
{
    float k_cos_theta = zm_dot3(a, b);
    if (approx_eql(k_cos_theta, 1.0f)) {
        orc_mat_identity(out);
        return;
    }
    float k = sqrtf(zm_dot3(a, a) * zm_dot3(b, b));
    if (approx_eql(k_cos_theta / k, -1.0f)) {
        const float ux[3] = {1.0f, 0.0f, 0.0f}, uy[3] = {0.0f, 1.0f, 0.0f};
        v3 o = vcross(vload(a), vload(ux));
        float ol[3] = {o.x, o.y, o.z};
        if (approx_eql(zm_dot3(ol, ol), 0.0f)) o = vcross(vload(a), vload(uy));
        o = zm_normalize3(o);
        float q[4] = {o.x, o.y, o.z, 0.0f};
        orc_mat_from_quat(q, out);
        return;
    }
    v3 v = vcross(vload(a), vload(b));
    float q[4] = {v.x, v.y, v.z, k + k_cos_theta};
    /* zmath normalize4: q / sqrt(dot4), dot4 = (z*z + w*w) + (x*x + y*y) */
    float len = sqrtf((q[2] * q[2] + q[3] * q[3]) + (q[0] * q[0] + q[1] * q[1]));
    for (int i = 0; i < 4; i++) q[i] = q[i] / len;
    orc_mat_from_quat(q, out);
}

/* aabb.zig:17-45 */
void orc_aabb_transform(const float m[16], const float mn[4], const float mx[4], float omn[4],
                        float omx[4])
{
    float p[8][4] = {
        {mn[0], mn[1], mn[2], mn[3]},
        {mx[0], mn[1], mn[2], 1.0f},
        {mn[0], mx[1], mn[2], 1.0f},
        {mn[0], mn[1], mx[2], 1.0f},
        {mn[0], mx[1], mx[2], 1.0f},
        {mx[0], mx[1], mn[2], 1.0f},
        {mx[0], mn[1], mx[2], 1.0f},
        {mx[0], mx[1], mx[2], mx[3]},
    };
    float q[4];
    orc_vec_mul_mat(p[0], m, q);
    float lo[4], hi[4];
    memcpy(lo, q, sizeof(q));
    memcpy(hi, q, sizeof(q));
    for (int i = 1; i < 8; i++) {
        orc_vec_mul_mat(p[i], m, q);
        for (int k = 0; k < 4; k++) {
            lo[k] = ominf(lo[k], q[k]);
            hi[k] = omaxf(hi[k], q[k]);
        }
    }
    memcpy(omn, lo, sizeof(lo));
    memcpy(omx, hi, sizeof(hi));
}

/* scene.zig:124-134 */
void orc_create_sphere(const float center[4], float radius, float transform[16], float aabb_min[4],
                       float aabb_max[4])
{
    float s[16], t[16];
    orc_mat_scaling(radius, radius, radius, s);
    orc_mat_translation(center[0], center[1], center[2], t);
    orc_mat_mul(s, t, transform);
    float rv[4] = {radius, radius, radius, 0.0f};
    for (int k = 0; k < 4; k++) {
        aabb_min[k] = center[k] - rv[k];
        aabb_max[k] = center[k] + rv[k];
    }
}

/* scene.zig:235-261 */
void orc_plane_mesh_transform(const float center[4], float side1, float side2, const float normal[4],
                              float transform[16])
{
    const float unit_y[4] = {0.0f, 1.0f, 0.0f, 0.0f};
    float rot[16], s[16], t[16], sr[16];
    orc_rotation_between_vectors(normal, unit_y, rot);
    orc_mat_scaling(side1, 1.0f, side2, s);
    orc_mat_mul(s, rot, sr);
    orc_mat_translation(center[0], center[1], center[2], t);
    orc_mat_mul(sr, t, transform);
}

/* scene.zig:274-287 */
void orc_mesh_aabb(const float *vertices, const uint32_t *vertex_indices, uint32_t vertex_indices_len,
                   const float transform[16], float ntmin[4], float ntmax[4], float aabb_min[4],
                   float aabb_max[4])
{
    float lo[4] = {INFINITY, INFINITY, INFINITY, 1.0f};
    float hi[4] = {-INFINITY, -INFINITY, -INFINITY, 1.0f};
    for (uint32_t t = 0; t < vertex_indices_len / 3u; t++) {
        const float *a = &vertices[4u * vertex_indices[3u * t]];
        const float *b = &vertices[4u * vertex_indices[3u * t + 1u]];
        const float *c = &vertices[4u * vertex_indices[3u * t + 2u]];
        for (int k = 0; k < 4; k++) {
            lo[k] = ominf(lo[k], ominf(ominf(a[k], b[k]), c[k]));
            hi[k] = omaxf(hi[k], omaxf(omaxf(a[k], b[k]), c[k]));
        }
    }
    memcpy(ntmin, lo, sizeof(lo));
    memcpy(ntmax, hi, sizeof(hi));
    orc_aabb_transform(transform, lo, hi, aabb_min, aabb_max);
}

/* ===================================================================== */
/* BVH build: bvh.zig:33-383                                               */
/* ===================================================================== */

/* Zig 0.11 std.rand.SplitMix64 */
uint64_t orc_splitmix64_next(uint64_t *state)
{
    *state += 0x9e3779b97f4a7c15ull;
    uint64_t z = *state;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

/* Zig 0.11 std.rand.Xoshiro256 (xoshiro256++) */
void orc_xoshiro_seed(uint64_t seed, uint64_t s[4])
{
    uint64_t g = seed;
    for (int i = 0; i < 4; i++) s[i] = orc_splitmix64_next(&g);
}

static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

uint64_t orc_xoshiro_next(uint64_t s[4])
{
    uint64_t r = rotl64(s[0] + s[3], 23) + s[0];
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return r;
}

/* Random.uintLessThan (Lemire with the PCG tweak), through uintAtMost */
uint64_t orc_xoshiro_range_at_most(uint64_t s[4], uint64_t at_most)
{
    uint64_t less_than = at_most + 1u;
    uint64_t x = orc_xoshiro_next(s);
    unsigned __int128 m = (unsigned __int128)x * less_than;
    uint64_t l = (uint64_t)m;
    if (l < less_than) {
        uint64_t t = (uint64_t)0 - less_than;
        if (t >= less_than) {
            t -= less_than;
            if (t >= less_than) t %= less_than;
        }
        while (l < t) {
            x = orc_xoshiro_next(s);
            m = (unsigned __int128)x * less_than;
            l = (uint64_t)m;
        }
    }
    return (uint64_t)(m >> 64);
}

/* Zig 0.11 std.sort.heap: generic over an index-based less/swap. */
typedef struct {
    int (*less)(void *ctx, uint32_t a, uint32_t b);
    void (*swap)(void *ctx, uint32_t a, uint32_t b);
    void *ctx;
} heap_ops;

static void sift_down(const heap_ops *op, uint32_t root, uint32_t n)
{
    uint32_t node = root;
    for (;;) {
        uint32_t child = 2u * node + 1u;
        if (child >= n) break;
        if (child + 1u < n && op->less(op->ctx, child, child + 1u)) child += 1u;
        if (!op->less(op->ctx, node, child)) break;
        op->swap(op->ctx, node, child);
        node = child;
    }
}

static void heap_sort(const heap_ops *op, uint32_t n)
{
    uint32_t i = n / 2u;
    while (i > 0u) {
        sift_down(op, i - 1u, n);
        i--;
    }
    i = n;
    while (i > 0u) {
        op->swap(op->ctx, 0u, i - 1u);
        sift_down(op, 0u, i - 1u);
        i--;
    }
}

typedef struct {
    float *keys;
    uint32_t *ids;
} key_ctx;

static int key_less(void *c, uint32_t a, uint32_t b)
{
    key_ctx *k = (key_ctx *)c;
    return k->keys[a] < k->keys[b];
}

static void key_swap(void *c, uint32_t a, uint32_t b)
{
    key_ctx *k = (key_ctx *)c;
    float t = k->keys[a];
    k->keys[a] = k->keys[b];
    k->keys[b] = t;
    uint32_t u = k->ids[a];
    k->ids[a] = k->ids[b];
    k->ids[b] = u;
}

void orc_heap_sort_keys(float *keys, uint32_t *ids, uint32_t n)
{
    key_ctx c = {keys, ids};
    heap_ops op = {key_less, key_swap, &c};
    heap_sort(&op, n);
}

/* growable arrays */
typedef struct {
    void *p;
    uint32_t len, cap;
    size_t elem;
} arr;

static void *arr_push(arr *a)
{
    if (a->len == a->cap) {
        a->cap = a->cap ? a->cap * 2u : 64u;
        a->p = realloc(a->p, a->cap * a->elem);
    }
    return (char *)a->p + (size_t)(a->len++) * a->elem;
}

typedef struct {
    float v0[4], v1[4], v2[4];
    uint32_t triangle_index;
    float mn[4], mx[4];
} bb_tri;

typedef struct {
    int kind; /* 0 sphere, 1 mesh, 2 instance */
    uint32_t idx;
} bb_leaf;

typedef struct {
    const orc_build_input *in;
    uint64_t prng[4];
    arr tlas, blas, normals, normal_indices, uvs, uv_indices, transforms, materials, textures;
    int32_t *material_index; /* per scene material, -1 until first use */
    uint32_t *mesh_bvh_id;
    int axis;
} bb_ctx;

static int tri_less(void *c, uint32_t a, uint32_t b)
{
    void **cc = (void **)c;
    bb_tri *t = (bb_tri *)cc[0];
    int axis = *(int *)cc[1];
    return t[a].mn[axis] < t[b].mn[axis];
}

static void tri_swap(void *c, uint32_t a, uint32_t b)
{
    void **cc = (void **)c;
    bb_tri *t = (bb_tri *)cc[0];
    bb_tri tmp = t[a];
    t[a] = t[b];
    t[b] = tmp;
}

static void leaf_box(const bb_ctx *b, bb_leaf l, const float **mn, const float **mx)
{
    if (l.kind == 0) {
        *mn = b->in->spheres[l.idx].aabb_min;
        *mx = b->in->spheres[l.idx].aabb_max;
    } else if (l.kind == 1) {
        *mn = b->in->meshes[l.idx].aabb_min;
        *mx = b->in->meshes[l.idx].aabb_max;
    } else {
        *mn = b->in->instances[l.idx].aabb_min;
        *mx = b->in->instances[l.idx].aabb_max;
    }
}

static int leaf_less(void *c, uint32_t a, uint32_t bb)
{
    void **cc = (void **)c;
    bb_ctx *b = (bb_ctx *)cc[0];
    bb_leaf *l = (bb_leaf *)cc[1];
    const float *amn, *amx, *bmn, *bmx;
    leaf_box(b, l[a], &amn, &amx);
    leaf_box(b, l[bb], &bmn, &bmx);
    return amn[b->axis] < bmn[b->axis];
}

static void leaf_swap(void *c, uint32_t a, uint32_t b)
{
    void **cc = (void **)c;
    bb_leaf *l = (bb_leaf *)cc[1];
    bb_leaf t = l[a];
    l[a] = l[b];
    l[b] = t;
}

static void set3(float *d, const float *s) { d[0] = s[0]; d[1] = s[1]; d[2] = s[2]; }

/* bvh.zig:221-270 */
static orc_bvh_node build_blas(bb_ctx *b, bb_tri *t, uint32_t n)
{
    orc_bvh_node node;
    memset(&node, 0, sizeof(node));
    if (n == 1u) {
        set3(node.lmin_or_v0, t[0].v0);
        set3(node.lmax_or_v1, t[0].v1);
        set3(node.rmin_or_v2, t[0].v2);
        node.left_or_custom_id = t[0].triangle_index;
        node.node_type = 3u;
        return node;
    }
    int axis = (int)orc_xoshiro_range_at_most(b->prng, 2);
    void *cc[2] = {t, &axis};
    heap_ops op = {tri_less, tri_swap, cc};
    heap_sort(&op, n);
    uint32_t mid = n / 2u;
    orc_bvh_node left = build_blas(b, t, mid);
    *(orc_bvh_node *)arr_push(&b->blas) = left;
    uint32_t left_id = b->blas.len - 1u;
    orc_bvh_node right = build_blas(b, t + mid, n - mid);
    *(orc_bvh_node *)arr_push(&b->blas) = right;
    uint32_t right_id = b->blas.len - 1u;
    float lmn[4] = {INFINITY, INFINITY, INFINITY, 1.0f}, lmx[4] = {-INFINITY, -INFINITY, -INFINITY, 1.0f};
    float rmn[4] = {INFINITY, INFINITY, INFINITY, 1.0f}, rmx[4] = {-INFINITY, -INFINITY, -INFINITY, 1.0f};
    for (uint32_t i = 0; i < mid; i++)
        for (int k = 0; k < 4; k++) {
            lmn[k] = ominf(lmn[k], t[i].mn[k]);
            lmx[k] = omaxf(lmx[k], t[i].mx[k]);
        }
    for (uint32_t i = mid; i < n; i++)
        for (int k = 0; k < 4; k++) {
            rmn[k] = ominf(rmn[k], t[i].mn[k]);
            rmx[k] = omaxf(rmx[k], t[i].mx[k]);
        }
    set3(node.lmin_or_v0, lmn);
    node.left_or_custom_id = left_id;
    set3(node.lmax_or_v1, lmx);
    node.right_or_material_index = right_id;
    set3(node.rmin_or_v2, rmn);
    node.node_type = 0u;
    set3(node.rmax_or_v3, rmx);
    return node;
}

/* bvh.zig:172-219 */
static void build_mesh(bb_ctx *b, uint32_t mi)
{
    const orc_build_mesh *m = &b->in->meshes[mi];
    uint32_t ntri = m->vertex_indices_len / 3u;
    bb_tri *t = (bb_tri *)malloc(sizeof(bb_tri) * (ntri ? ntri : 1u));
    uint32_t base = b->normal_indices.len / 3u;
    for (uint32_t i = 0; i < ntri; i++) {
        memcpy(t[i].v0, &m->vertices[4u * m->vertex_indices[3u * i]], 16);
        memcpy(t[i].v1, &m->vertices[4u * m->vertex_indices[3u * i + 1u]], 16);
        memcpy(t[i].v2, &m->vertices[4u * m->vertex_indices[3u * i + 2u]], 16);
        for (int k = 0; k < 4; k++) {
            t[i].mn[k] = ominf(ominf(t[i].v0[k], t[i].v1[k]), t[i].v2[k]);
            t[i].mx[k] = omaxf(omaxf(t[i].v0[k], t[i].v1[k]), t[i].v2[k]);
        }
        t[i].triangle_index = base + i;
    }
    uint32_t nbase = b->normals.len;
    for (uint32_t i = 0; i < m->normal_indices_len; i++)
        *(uint32_t *)arr_push(&b->normal_indices) = m->normal_indices[i] + nbase;
    for (uint32_t i = 0; i < m->normals_len; i++) memcpy(arr_push(&b->normals), &m->normals[4u * i], 16);
    uint32_t ubase = b->uvs.len;
    for (uint32_t i = 0; i < m->uv_indices_len; i++)
        *(uint32_t *)arr_push(&b->uv_indices) = m->uv_indices[i] + ubase;
    for (uint32_t i = 0; i < m->uvs_len; i++) memcpy(arr_push(&b->uvs), &m->uvs[2u * i], 8);
    orc_bvh_node root = build_blas(b, t, ntri);
    *(orc_bvh_node *)arr_push(&b->blas) = root;
    b->mesh_bvh_id[mi] = b->blas.len - 1u;
    free(t);
}

static void append_transform(bb_ctx *b, const float m[16])
{
    float inv[16];
    orc_mat_inverse(m, inv);
    memcpy(arr_push(&b->transforms), inv, 64);
    memcpy(arr_push(&b->transforms), m, 64);
}

/* bvh.zig:369-383 */
static uint32_t material_index(bb_ctx *b, uint32_t mi)
{
    if (b->material_index[mi] >= 0) return (uint32_t)b->material_index[mi];
    const orc_build_material *m = &b->in->materials[mi];
    orc_material g;
    memset(&g, 0, sizeof(g));
    g.albedo[0] = 1.0f;
    g.albedo[1] = 0.0f;
    g.albedo[2] = 1.0f;
    g.albedo_texture_index = 0xffffffffu;
    if (m->texture >= 0) {
        *(uint32_t *)arr_push(&b->textures) = (uint32_t)m->texture;
        g.albedo_texture_index = b->textures.len - 1u;
    } else {
        set3(g.albedo, m->albedo);
    }
    g.fuzz = m->fuzz;
    g.ior = m->ior;
    g.type = m->type;
    *(orc_material *)arr_push(&b->materials) = g;
    b->material_index[mi] = (int32_t)(b->materials.len - 1u);
    return b->materials.len - 1u;
}

/* bvh.zig:272-362 */
static orc_bvh_node build_tlas(bb_ctx *b, bb_leaf *l, uint32_t n)
{
    orc_bvh_node node;
    memset(&node, 0, sizeof(node));
    if (n == 1u) {
        const float *mn, *mx;
        leaf_box(b, l[0], &mn, &mx);
        set3(node.lmin_or_v0, mn);
        set3(node.lmax_or_v1, mx);
        if (l[0].kind == 0) {
            const orc_build_sphere *s = &b->in->spheres[l[0].idx];
            append_transform(b, s->transform);
            node.transform_id = b->transforms.len / 2u - 1u;
            node.right_or_material_index = material_index(b, s->material);
            node.node_type = 1u;
        } else if (l[0].kind == 1) {
            const orc_build_mesh *m = &b->in->meshes[l[0].idx];
            append_transform(b, m->transform);
            node.transform_id = b->transforms.len / 2u - 1u;
            node.left_or_custom_id = b->mesh_bvh_id[l[0].idx];
            node.right_or_material_index = material_index(b, m->material);
            node.node_type = 2u;
        } else {
            const orc_build_instance *mi = &b->in->instances[l[0].idx];
            append_transform(b, mi->transform);
            node.transform_id = b->transforms.len / 2u - 1u;
            node.right_or_material_index = material_index(b, mi->material);
            node.node_type = 2u;
            node.left_or_custom_id = b->mesh_bvh_id[mi->mesh];
        }
        return node;
    }
    b->axis = (int)orc_xoshiro_range_at_most(b->prng, 2);
    void *cc[2] = {b, l};
    heap_ops op = {leaf_less, leaf_swap, cc};
    heap_sort(&op, n);
    uint32_t mid = n / 2u;
    orc_bvh_node left = build_tlas(b, l, mid);
    *(orc_bvh_node *)arr_push(&b->tlas) = left;
    uint32_t left_id = b->tlas.len - 1u;
    float lmn[4] = {INFINITY, INFINITY, INFINITY, 1.0f}, lmx[4] = {-INFINITY, -INFINITY, -INFINITY, 1.0f};
    for (uint32_t i = 0; i < mid; i++) {
        const float *mn, *mx;
        leaf_box(b, l[i], &mn, &mx);
        for (int k = 0; k < 4; k++) {
            lmn[k] = ominf(lmn[k], mn[k]);
            lmx[k] = omaxf(lmx[k], mx[k]);
        }
    }
    orc_bvh_node right = build_tlas(b, l + mid, n - mid);
    *(orc_bvh_node *)arr_push(&b->tlas) = right;
    uint32_t right_id = b->tlas.len - 1u;
    float rmn[4] = {INFINITY, INFINITY, INFINITY, 1.0f}, rmx[4] = {-INFINITY, -INFINITY, -INFINITY, 1.0f};
    for (uint32_t i = mid; i < n; i++) {
        const float *mn, *mx;
        leaf_box(b, l[i], &mn, &mx);
        for (int k = 0; k < 4; k++) {
            rmn[k] = ominf(rmn[k], mn[k]);
            rmx[k] = omaxf(rmx[k], mx[k]);
        }
    }
    set3(node.lmin_or_v0, lmn);
    node.left_or_custom_id = left_id;
    set3(node.lmax_or_v1, lmx);
    node.right_or_material_index = right_id;
    set3(node.rmin_or_v2, rmn);
    node.node_type = 0u;
    set3(node.rmax_or_v3, rmx);
    return node;
}

int orc_bvh_build(const orc_build_input *in, orc_build_output *out)
{
    uint64_t prng[4];
    orc_xoshiro_seed(1244u, prng); /* bvh.zig:12 DefaultPrng.init(1244) */
    return orc_bvh_build_stream(in, out, prng);
}

int orc_bvh_build_stream(const orc_build_input *in, orc_build_output *out, uint64_t prng[4])
{
    memset(out, 0, sizeof(*out));
    uint32_t shapes = in->spheres_len + in->meshes_len + in->instances_len;
    if (shapes == 0u) return -1; /* bvh.zig:35-37 panics on an empty scene */
    for (uint32_t i = 0; i < in->meshes_len; i++)
        if (in->meshes[i].vertex_indices_len < 3u) return -2; /* bvh.zig:222-223 */
    bb_ctx b;
    memset(&b, 0, sizeof(b));
    b.in = in;
    memcpy(b.prng, prng, sizeof(b.prng)); /* the process's generator (bvh.zig:12-13) */
    b.tlas.elem = b.blas.elem = sizeof(orc_bvh_node);
    b.normals.elem = 16;
    b.normal_indices.elem = b.uv_indices.elem = b.textures.elem = 4;
    b.uvs.elem = 8;
    b.transforms.elem = 64;
    b.materials.elem = sizeof(orc_material);
    b.material_index = (int32_t *)malloc(sizeof(int32_t) * (in->materials_len + 1u));
    for (uint32_t i = 0; i < in->materials_len; i++) b.material_index[i] = -1;
    b.mesh_bvh_id = (uint32_t *)calloc(in->meshes_len + 1u, sizeof(uint32_t));
    bb_leaf *leaves = (bb_leaf *)malloc(sizeof(bb_leaf) * shapes);
    uint32_t nl = 0;
    /* bvh.zig:99-105: spheres, then mesh instances, then meshes */
    for (uint32_t i = 0; i < in->spheres_len; i++) { leaves[nl].kind = 0; leaves[nl++].idx = i; }
    for (uint32_t i = 0; i < in->instances_len; i++) { leaves[nl].kind = 2; leaves[nl++].idx = i; }
    for (uint32_t i = 0; i < in->meshes_len; i++) {
        leaves[nl].kind = 1;
        leaves[nl++].idx = i;
        build_mesh(&b, i);
    }
    orc_bvh_node root = build_tlas(&b, leaves, nl);
    *(orc_bvh_node *)arr_push(&b.tlas) = root;
    free(leaves);
    free(b.material_index);
    free(b.mesh_bvh_id);
    out->tlas = (orc_bvh_node *)b.tlas.p;
    out->tlas_len = b.tlas.len;
    out->blas = (orc_bvh_node *)b.blas.p;
    out->blas_len = b.blas.len;
    out->normals = (float *)b.normals.p;
    out->normals_len = b.normals.len;
    out->normal_indices = (uint32_t *)b.normal_indices.p;
    out->normal_indices_len = b.normal_indices.len;
    out->uvs = (float *)b.uvs.p;
    out->uvs_len = b.uvs.len;
    out->uv_indices = (uint32_t *)b.uv_indices.p;
    out->uv_indices_len = b.uv_indices.len;
    out->transforms = (float *)b.transforms.p;
    out->transforms_len = b.transforms.len;
    out->materials = (orc_material *)b.materials.p;
    out->materials_len = b.materials.len;
    out->textures = (uint32_t *)b.textures.p;
    out->textures_len = b.textures.len;
    memcpy(prng, b.prng, sizeof(b.prng)); /* the next build continues the stream */
    return 0;
}

void orc_bvh_free(orc_build_output *out)
{
    free(out->tlas);
    free(out->blas);
    free(out->normals);
    free(out->normal_indices);
    free(out->uvs);
    free(out->uv_indices);
    free(out->transforms);
    free(out->materials);
    free(out->textures);
    memset(out, 0, sizeof(*out));
}
