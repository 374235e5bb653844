// orn_common.h -- FP32 vector helpers, the per-pixel RNG and the FP32
// transcendental functions used by the gfx950 kernels.
//
// Every expression is evaluated as written (the library is compiled with
// -ffp-contract=off); the only fused multiply-adds are explicit fmaf calls.
// That makes the kernels' arithmetic reproducible on any IEEE-754 machine.
//
// WGSL leaves the precision of pow/sin/cos/acos/atan2 to the implementation
// (the reference runs its megakernel on whatever the adapter provides), so
// the kernels define them here:
//   sin, cos     Cody-Waite reduction by pi/2 with a two-float pi/2 and FMA,
//                Taylor polynomials of degree 9 / 10 on [-pi/4, pi/4]
//   pow(x, y)    exp2(y * log2(x)), the formula the WGSL spec's error bound is
//                stated for; log2 by the atanh series on m in [sqrt(1/2), sqrt(2)],
//                exp2 by a degree-7 polynomial in the fraction
//   acos, atan2  the asinf / atanf polynomials of the Cephes library
//                (S. L. Moshier), with the usual range reductions
#pragma once

#include <stdint.h>

#include <hip/hip_runtime.h>

#define ORN_HD __host__ __device__ __forceinline__

namespace orn {

struct f3 {
    float x, y, z;
};

ORN_HD f3 mk(float x, float y, float z) { return f3{x, y, z}; }
ORN_HD f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
ORN_HD f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
ORN_HD f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
ORN_HD f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
ORN_HD f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
ORN_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
ORN_HD f3 cross(f3 a, f3 b)
{
    return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// normalize(v) = v * (1 / sqrt(dot(v, v))), correctly rounded sqrt and divide
ORN_HD f3 normalize(f3 a)
{
    float inv = 1.0f / sqrtf(dot(a, a));
    return f3{a.x * inv, a.y * inv, a.z * inv};
}
ORN_HD f3 xyz(float4 v) { return f3{v.x, v.y, v.z}; }

// Mask of the wave's active lanes for which b holds. Taking the bool directly
// lets a comparison feed the mask; HIP's __ballot(int) first widens it to a
// lane value and compares that again.
__device__ __forceinline__ uint64_t lanes_where(bool b) { return __builtin_amdgcn_ballot_w64(b); }

ORN_HD uint32_t as_u32(float f) { return __builtin_bit_cast(uint32_t, f); }
ORN_HD float as_f32(uint32_t u) { return __builtin_bit_cast(float, u); }

constexpr float kPi = 3.1415927410125732f;    // utility.wgsl:1 rounded to f32
constexpr float kHalfPi = 1.5707963705062866f;
constexpr float kQuarterPi = 0.7853981852531433f;

// ---------------------------------------------------------------------------
// Per-pixel RNG, random.wgsl:11-22. The "old_state" step adds the two PCG
// constants instead of multiplying; 747796405 + 2891336453 wraps to 3639132858.
// ---------------------------------------------------------------------------
ORN_HD uint32_t rng_next(uint32_t &s)
{
    uint32_t x = s + 3639132858u;
    uint32_t w = ((x >> ((x >> 28u) + 4u)) ^ x) * 277803737u;
    s = (w >> 22u) ^ w;
    return s;
}

// f32(u32) rounds to nearest even, so a draw can be exactly 1.0
ORN_HD float rng_f32(uint32_t &s) { return static_cast<float>(rng_next(s)) * 2.3283064365386963e-10f; }

// ---------------------------------------------------------------------------
// sin / cos
// ---------------------------------------------------------------------------
ORN_HD void sin_cos(float x, float &s, float &c)
{
    const float k = floorf(x * 0.6366197466850281f + 0.5f);
    float r = fmaf(k, -1.5707963705062866f, x);
    r = fmaf(k, 4.371138828673793e-08f, r);
    const int q = static_cast<int>(k) & 3;
    const float r2 = r * r;
    float ps = fmaf(r2, 2.7557318844628753e-06f, -0.00019841270113829523f);
    ps = fmaf(r2, ps, 0.008333333767950535f);
    ps = fmaf(r2, ps, -0.1666666716337204f);
    const float sn = fmaf(r * r2, ps, r);
    float pc = fmaf(r2, -2.755731998149713e-07f, 2.4801587642286904e-05f);
    pc = fmaf(r2, pc, -0.0013888889225199819f);
    pc = fmaf(r2, pc, 0.0416666679084301f);
    pc = fmaf(r2, pc, -0.5f);
    const float cs = fmaf(r2, pc, 1.0f);
    switch (q) {
    case 0: s = sn; c = cs; break;
    case 1: s = cs; c = -sn; break;
    case 2: s = -sn; c = -cs; break;
    default: s = -cs; c = sn; break;
    }
}

// ---------------------------------------------------------------------------
// pow(x, y) = exp2(y * log2(x))
// ---------------------------------------------------------------------------
ORN_HD float log2_positive(float x)
{
    uint32_t bits = as_u32(x);
    int e = 0;
    if (bits < 0x00800000u) {  // subnormal input: rescale by 2^24
        x = x * 16777216.0f;
        bits = as_u32(x);
        e = -24;
    }
    e += static_cast<int>((bits >> 23) & 0xffu) - 127;
    float m = as_f32((bits & 0x007fffffu) | 0x3f800000u);
    if (m > 1.4142135381698608f) {
        m = m * 0.5f;
        e += 1;
    }
    const float s = (m - 1.0f) / (m + 1.0f);  // ln m = 2 atanh(s)
    const float s2 = s * s;
    float p = fmaf(s2, 0.09090909361839294f, 0.1111111119389534f);
    p = fmaf(s2, p, 0.1428571492433548f);
    p = fmaf(s2, p, 0.20000000298023224f);
    p = fmaf(s2, p, 0.3333333432674408f);
    p = fmaf(s2, p, 1.0f);
    const float ln_m = 2.0f * s * p;
    return fmaf(ln_m, 1.4426950216293335f, static_cast<float>(e));
}

ORN_HD float exp2_approx(float z)
{
    if (z != z) return z;
    if (z >= 128.0f) return __builtin_inff();
    if (z < -150.0f) return 0.0f;
    const float n = floorf(z + 0.5f);
    const float f = z - n;  // exact, |f| <= 1/2
    float p = fmaf(f, 1.5252733646775596e-05f, 0.0001540352968731895f);
    p = fmaf(f, p, 0.0013333557872101665f);
    p = fmaf(f, p, 0.009618128649890423f);
    p = fmaf(f, p, 0.05550410971045494f);
    p = fmaf(f, p, 0.24022650718688965f);
    p = fmaf(f, p, 0.6931471824645996f);
    p = fmaf(f, p, 1.0f);
    const int ni = static_cast<int>(n);
    if (ni >= -126) {
        if (ni > 127) return __builtin_inff();
        return p * as_f32(static_cast<uint32_t>(ni + 127) << 23);
    }
    // subnormal result: two exact power-of-two scalings
    const float t = p * as_f32(static_cast<uint32_t>(ni + 64 + 127) << 23);
    return t * as_f32(static_cast<uint32_t>(127 - 64) << 23);
}

ORN_HD float pow_f(float x, float y)
{
    if (x != x || y != y) return x + y;
    if (x < 0.0f) return __builtin_nanf("");
    if (x == 0.0f) return (y > 0.0f) ? 0.0f : ((y == 0.0f) ? 1.0f : __builtin_inff());
    if (x == __builtin_inff()) return (y > 0.0f) ? __builtin_inff() : ((y == 0.0f) ? 1.0f : 0.0f);
    return exp2_approx(y * log2_positive(x));
}

// ---------------------------------------------------------------------------
// acos / atan2 (Cephes polynomials)
// ---------------------------------------------------------------------------
ORN_HD float asin_poly(float t, float z)
{
    float p = fmaf(z, 4.2163199048e-2f, 2.4181311049e-2f);
    p = fmaf(z, p, 4.5470025998e-2f);
    p = fmaf(z, p, 7.4953002686e-2f);
    p = fmaf(z, p, 1.6666752422e-1f);
    return fmaf(p * z, t, t);
}

ORN_HD float acos_f(float x)
{
    const float a = fabsf(x);
    if (a > 0.5f) {
        const float z = 0.5f * (1.0f - a);
        const float r = 2.0f * asin_poly(sqrtf(z), z);
        return (x < 0.0f) ? (kPi - r) : r;
    }
    return kHalfPi - asin_poly(x, x * x);
}

ORN_HD float atan2_f(float y, float x)
{
    if (x != x || y != y) return x + y;
    const float ax = fabsf(x), ay = fabsf(y);
    const float hi = ax > ay ? ax : ay;
    const float lo = ax > ay ? ay : ax;
    float a = 0.0f;
    if (hi != 0.0f) {
        float t = lo / hi;
        float base = 0.0f;
        if (t > 0.4142135679721832f) {  // tan(pi/8)
            base = kQuarterPi;
            t = (t - 1.0f) / (t + 1.0f);
        }
        const float z = t * t;
        float p = fmaf(z, 8.05374449538e-2f, -1.38776856032e-1f);
        p = fmaf(z, p, 1.99777106478e-1f);
        p = fmaf(z, p, -3.33329491539e-1f);
        a = base + fmaf(p * z, t, t);
    }
    if (ay > ax) a = kHalfPi - a;
    if (x < 0.0f || (x == 0.0f && (as_u32(x) >> 31))) a = kPi - a;
    if (as_u32(y) >> 31) a = -a;
    return a;
}

// ---------------------------------------------------------------------------
// Sampling helpers, random.wgsl:36-87
// ---------------------------------------------------------------------------

// theta is uniform in [0, pi]: this is the reference's (non-uniform) sphere
ORN_HD f3 random_in_unit_sphere(uint32_t &s)
{
    const float r = pow_f(rng_f32(s), 0.33333f);
    const float theta = kPi * rng_f32(s);
    const float phi = (2.0f * kPi) * rng_f32(s);
    float st, ct, sp, cp;
    sin_cos(theta, st, ct);
    sin_cos(phi, sp, cp);
    return f3{r * st * cp, r * st * sp, r * ct};
}

ORN_HD f3 random_unit_vector(uint32_t &s) { return normalize(random_in_unit_sphere(s)); }

ORN_HD f3 random_in_unit_disk(uint32_t &s)
{
    const float r = sqrtf(rng_f32(s));
    const float alpha = (2.0f * kPi) * rng_f32(s);
    float sa, ca;
    sin_cos(alpha, sa, ca);
    return f3{r * ca, r * sa, 0.0f};
}

}  // namespace orn
