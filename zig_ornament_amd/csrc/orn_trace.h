// orn_trace.h -- device-side scene view, BVH traversal and shading for the
// gfx950 path-tracing kernels.
//
// Semantics follow the reference's WGSL megakernel
// (src/wgpu_backend/shaders/*.wgsl) operation for operation, so that the
// closest hit, every RNG draw and every accumulated value are reproduced
// bit for bit. The layout work is the MI355X part: nodes are fetched as four
// 16-byte vector loads, the traversal stack lives in LDS (one column per
// lane, stride = block size, so the 64 lanes of a wave hit 64 different
// banks), and materials are two 16-byte loads.
#pragma once

#include "orn_common.h"

namespace orn {

constexpr uint32_t kFinishedTraverseBlas = 0xffffffffu;  // pathtracer.wgsl:19
constexpr float kTMaxInit = 3.40282e+38f;                 // bvh.wgsl:23

// One texture in the texel pool. format: components | (hdr << 8)
struct DevTexture {
    uint64_t offset;  // byte offset in the texel pool
    uint32_t width;
    uint32_t height;
    uint32_t components;
    uint32_t hdr;
};

// BVH4 child references: a BVH4 node index, a BVH2 leaf (kLeafRef | index
// into the BVH2 array of the same level) or an empty slot.
constexpr uint32_t kLeafRef = 0x80000000u;
constexpr uint32_t kEmptyRef = 0xfffffffeu;

// The bound buffers of pathtracer.wgsl:7-17, resident in HBM, plus the
// derived BVH4 arrays (8 float4 = 128 B per node: children's min x, y, z,
// max x, y, z as float4 lanes, then the 4 child references).
struct DevScene {
    const float4 *__restrict__ tlas;        // 4 float4 per node
    const float4 *__restrict__ blas;        // 4 float4 per node
    const float4 *__restrict__ tlas4;       // BVH4 of the TLAS
    const float4 *__restrict__ blas4;       // BVH4 of every BLAS
    const float4 *__restrict__ tlas_leaves; // TLAS copy whose mesh leaves point at BLAS BVH4 roots
    const float4 *__restrict__ tlas2;       // TLAS/BLAS copies whose child indices (and mesh
    const float4 *__restrict__ blas2;       //   leaves' BLAS roots) carry kLeafRef for leaves
    const float4 *__restrict__ normals;     // 1 float4 per normal
    const uint32_t *__restrict__ normal_indices;
    const float2 *__restrict__ uvs;
    const uint32_t *__restrict__ uv_indices;
    const float4 *__restrict__ transforms;  // 4 float4 (columns) per matrix
    const uint4 *__restrict__ materials;    // 2 uint4 per material
    const DevTexture *__restrict__ textures;
    const uint8_t *__restrict__ texels;
    uint32_t tlas_len;
    uint32_t textures_len;
    uint32_t stack_slots;  // LDS stack entries per lane
    uint32_t tlas4_root;   // reference of the TLAS root in the BVH4
    uint32_t tlas2_root;   // index of the TLAS root, kLeafRef set when it is a leaf
    uint32_t pad[3];
};

struct Hit {
    float t;
    uint32_t material;
    uint32_t node_type;
    uint32_t inv_tid;  // 2 * transform_id
    uint32_t tri_id;   // 3 * triangle index
    float u, v;
};

struct Ray {
    f3 o, d;
};

// bvh.wgsl:139-174
__device__ __forceinline__ float copysign_ref(float a, float b) { return (b < 0.0f) ? -a : a; }

__device__ __forceinline__ f3 safe_invdir(f3 d)
{
    const float eps = 1e-5f;
    const float x = (fabsf(d.x) > eps) ? d.x : copysign_ref(eps, d.x);
    const float y = (fabsf(d.y) > eps) ? d.y : copysign_ref(eps, d.y);
    const float z = (fabsf(d.z) > eps) ? d.z : copysign_ref(eps, d.z);
    return f3{1.0f / x, 1.0f / y, 1.0f / z};
}

// bvh.wgsl:184-192: slab test with fused multiply-adds
__device__ __forceinline__ bool aabb_hit(float3 mn, float3 mx, f3 inv, f3 ox, float t_min,
                                         float t_max)
{
    const float fx = fmaf(mx.x, inv.x, ox.x), fy = fmaf(mx.y, inv.y, ox.y),
                fz = fmaf(mx.z, inv.z, ox.z);
    const float nx = fmaf(mn.x, inv.x, ox.x), ny = fmaf(mn.y, inv.y, ox.y),
                nz = fmaf(mn.z, inv.z, ox.z);
    const float far_t = fminf(fminf(fminf(fmaxf(fx, nx), fmaxf(fy, ny)), fmaxf(fz, nz)), t_max);
    const float near_t = fmaxf(fmaxf(fmaxf(fminf(fx, nx), fminf(fy, ny)), fminf(fz, nz)), t_min);
    return near_t <= far_t;
}

// bvh.wgsl:194-226: Moller-Trumbore as the reference writes it (no
// determinant test; NaN falls through the comparisons)
__device__ __forceinline__ float triangle_hit(const Ray &r, f3 v1, f3 v2, f3 v3, float t_min,
                                              float t_max, float &hu, float &hv)
{
    const f3 e1 = v2 - v1;
    const f3 e2 = v3 - v1;
    const f3 s1 = cross(r.d, e2);
    const float invd = 1.0f / dot(s1, e1);
    const f3 d = r.o - v1;
    const float u = dot(d, s1) * invd;
    if (u < 0.0f || u > 1.0f) return t_max;
    const f3 s2 = cross(d, e1);
    const float v = dot(r.d, s2) * invd;
    if (v < 0.0f || u + v > 1.0f) return t_max;
    const float t = dot(e2, s2) * invd;
    if (t < t_min || t > t_max) return t_max;
    hu = u;
    hv = v;
    return t;
}

// bvh.wgsl:228-249: unit sphere at the origin of object space
__device__ __forceinline__ float sphere_hit(const Ray &r, float t_min, float t_max)
{
    const float a = r.d.x * r.d.x + r.d.y * r.d.y + r.d.z * r.d.z;
    const float half_b = dot(r.o, r.d);
    const float c = (r.o.x * r.o.x + r.o.y * r.o.y + r.o.z * r.o.z) - 1.0f;
    const float disc = half_b * half_b - a * c;
    if (disc < 0.0f) return t_max;
    const float sq = sqrtf(disc);
    float t = (-half_b - sq) / a;
    if (t < t_min || t_max < t) {
        t = (-half_b + sq) / a;
        if (t < t_min || t_max < t) return t_max;
    }
    return t;
}

// transform.wgsl: the 16 floats are read column-major; c0..c3 are columns
__device__ __forceinline__ f3 xform_point(const float4 *m, f3 p)
{
    const float4 c0 = m[0], c1 = m[1], c2 = m[2], c3 = m[3];
    return f3{c0.x * p.x + c1.x * p.y + c2.x * p.z + c3.x * 1.0f,
              c0.y * p.x + c1.y * p.y + c2.y * p.z + c3.y * 1.0f,
              c0.z * p.x + c1.z * p.y + c2.z * p.z + c3.z * 1.0f};
}

__device__ __forceinline__ Ray xform_ray(const float4 *m, const Ray &r)
{
    const float4 c0 = m[0], c1 = m[1], c2 = m[2], c3 = m[3];
    Ray o;
    o.o = f3{c0.x * r.o.x + c1.x * r.o.y + c2.x * r.o.z + c3.x * 1.0f,
             c0.y * r.o.x + c1.y * r.o.y + c2.y * r.o.z + c3.y * 1.0f,
             c0.z * r.o.x + c1.z * r.o.y + c2.z * r.o.z + c3.z * 1.0f};
    o.d = f3{c0.x * r.d.x + c1.x * r.d.y + c2.x * r.d.z + c3.x * 0.0f,
             c0.y * r.d.x + c1.y * r.d.y + c2.y * r.d.z + c3.y * 0.0f,
             c0.z * r.d.x + c1.z * r.d.y + c2.z * r.d.z + c3.z * 0.0f};
    return o;
}

// transform.wgsl:16-19: transpose(M) * (n, 0), i.e. dot(column_i, (n, 0))
__device__ __forceinline__ f3 xform_normal(const float4 *m, f3 n)
{
    const float4 c0 = m[0], c1 = m[1], c2 = m[2];
    return f3{c0.x * n.x + c0.y * n.y + c0.z * n.z + c0.w * 0.0f,
              c1.x * n.x + c1.y * n.y + c1.z * n.z + c1.w * 0.0f,
              c2.x * n.x + c2.y * n.y + c2.z * n.z + c2.w * 0.0f};
}

// Scene content known at launch: kernels are instantiated without the code
// (and registers) of primitive kinds a scene does not contain.
enum SceneFeatures : uint32_t {
    kHasSpheres = 1u,
    kHasMeshes = 2u,
    kHasTextures = 4u,
};

// bvh.wgsl:14-137 as a resumable state machine: trav_begin() sets up one ray,
// trav_step() visits one node and returns false once the ray is finished.
// Same visiting order as the reference (left pushed before right, so the
// right child is taken first) and the same strict `t < t_max` update, so
// ties between equal distances resolve the same way. The node taken next is
// kept in a register: only nodes the reference leaves on its stack go to the
// lane's LDS column (`stack`, stride = block size).
struct Traversal {
    Ray ray;        // current space (world, or object space inside a BLAS)
    f3 inv, ox;     // safe_invdir(ray.d) and -ray.o * inv
    f3 winv, wox;   // the same of the world-space ray, restored on leaving a BLAS
    float t_max;
    uint32_t addr;  // next node to visit
    uint32_t top;   // pending entries in the LDS column
    uint32_t material_index, inv_tid;  // of the mesh being traversed
    bool in_tlas;
    bool hit;
    Hit h;
};

// `root`: the TLAS root reference in the traversed format (BVH2 or BVH4).
__device__ __forceinline__ void trav_begin(uint32_t root, const Ray &world, Traversal &tr)
{
    tr.ray = world;
    tr.inv = safe_invdir(world.d);
    tr.ox = (-world.o) * tr.inv;
    tr.winv = tr.inv;
    tr.wox = tr.ox;
    tr.t_max = kTMaxInit;
    tr.addr = root;
    tr.top = 0u;
    tr.material_index = 0u;
    tr.inv_tid = 0u;
    tr.in_tlas = true;
    tr.hit = false;
}

// An empty asm statement that reads and "writes" one register of each of a
// node's four 16-byte loads: the compiler must have all four loaded at this
// point and cannot sink any of them into a branch. No instruction is emitted.
__device__ __forceinline__ void pin_node(float4 &a, float4 &b, float4 &c, float4 &d)
{
    asm volatile("" : "+v"(a.x), "+v"(b.x), "+v"(c.x), "+v"(d.x));
}

// Takes the next pending node off the lane's stack; false when none is left.
template <uint32_t F>
__device__ __forceinline__ bool trav_pop(const DevScene &sc, const Ray &world, uint32_t *stack,
                                         uint32_t stride, Traversal &tr)
{
    if (tr.top == 0u) return false;
    tr.addr = stack[(--tr.top) * stride];
    if ((F & kHasMeshes) && tr.addr == kFinishedTraverseBlas) {
        // back in the TLAS: the world-space ray and its reciprocal, kept
        // from trav_begin
        tr.in_tlas = true;
        tr.ray = world;
        tr.inv = tr.winv;
        tr.ox = tr.wox;
        if (tr.top == 0u) return false;
        tr.addr = stack[(--tr.top) * stride];
    }
    return true;
}

// One BVH4 node: the hit children are pushed in slot order and the last hit
// one is taken next, so they come off the stack last slot first -- the
// reference's right-before-left order, two levels at a time.
__device__ __forceinline__ bool bvh4_node(const float4 *nd, uint32_t *stack, uint32_t stride,
                                          uint32_t cap, float t_min, Traversal &tr)
{
    const float4 lx = nd[0], ly = nd[1], lz = nd[2], hx = nd[3], hy = nd[4], hz = nd[5];
    const uint4 c = *reinterpret_cast<const uint4 *>(nd + 6);
    const bool h0 = c.x != kEmptyRef &&
                    aabb_hit(make_float3(lx.x, ly.x, lz.x), make_float3(hx.x, hy.x, hz.x), tr.inv, tr.ox, t_min, tr.t_max);
    const bool h1 = c.y != kEmptyRef &&
                    aabb_hit(make_float3(lx.y, ly.y, lz.y), make_float3(hx.y, hy.y, hz.y), tr.inv, tr.ox, t_min, tr.t_max);
    const bool h2 = c.z != kEmptyRef &&
                    aabb_hit(make_float3(lx.z, ly.z, lz.z), make_float3(hx.z, hy.z, hz.z), tr.inv, tr.ox, t_min, tr.t_max);
    const bool h3 = c.w != kEmptyRef &&
                    aabb_hit(make_float3(lx.w, ly.w, lz.w), make_float3(hx.w, hy.w, hz.w), tr.inv, tr.ox, t_min, tr.t_max);
    uint32_t next = kEmptyRef;
    if (h0) next = c.x;
    if (h1) {
        if (next != kEmptyRef && tr.top < cap) stack[(tr.top++) * stride] = next;
        next = c.y;
    }
    if (h2) {
        if (next != kEmptyRef && tr.top < cap) stack[(tr.top++) * stride] = next;
        next = c.z;
    }
    if (h3) {
        if (next != kEmptyRef && tr.top < cap) stack[(tr.top++) * stride] = next;
        next = c.w;
    }
    if (next == kEmptyRef) return false;
    tr.addr = next;
    return true;
}

// Node references carry kLeafRef when they name a leaf (in the BVH4 by
// construction, in the BVH2 through the flagged copies), so a lane knows
// whether its next step is an internal node or a leaf before loading it.
// trav_internal() visits an internal node, trav_leaf() a leaf; both return
// false once the ray is finished.
template <uint32_t F, uint32_t W>
__device__ __forceinline__ bool trav_internal(const DevScene &sc, float t_min, const Ray &world,
                                              uint32_t *stack, uint32_t stride, Traversal &tr)
{
    const uint32_t cap = sc.stack_slots;
    const bool blas = (F & kHasMeshes) && !tr.in_tlas;
    if (W == 4u) {
        const float4 *nd = (blas ? sc.blas4 : sc.tlas4) + 8u * tr.addr;
        if (bvh4_node(nd, stack, stride, cap, t_min, tr)) return true;
        return trav_pop<F>(sc, world, stack, stride, tr);
    }
    const float4 *node = (blas ? sc.blas2 : sc.tlas2) + 4u * tr.addr;
    float4 n0 = node[0], n1 = node[1], n2 = node[2], n3 = node[3];
    const bool l = aabb_hit(make_float3(n0.x, n0.y, n0.z), make_float3(n1.x, n1.y, n1.z), tr.inv,
                            tr.ox, t_min, tr.t_max);
    const bool r = aabb_hit(make_float3(n2.x, n2.y, n2.z), make_float3(n3.x, n3.y, n3.z), tr.inv,
                            tr.ox, t_min, tr.t_max);
    // both hit: the left child waits on the stack and the right one is next
    if (l && r && tr.top < cap) stack[(tr.top++) * stride] = as_u32(n0.w);
    if (l || r) {
        tr.addr = r ? as_u32(n1.w) : as_u32(n0.w);
        return true;
    }
    return trav_pop<F>(sc, world, stack, stride, tr);
}

template <uint32_t F, uint32_t W>
__device__ __forceinline__ bool trav_leaf(const DevScene &sc, float t_min, const Ray &world,
                                          uint32_t *stack, uint32_t stride, Traversal &tr)
{
    const uint32_t cap = sc.stack_slots;
    const uint32_t idx = tr.addr & ~kLeafRef;
    const float4 *base = (F & kHasMeshes) && !tr.in_tlas ? sc.blas2 : (W == 4u ? sc.tlas_leaves : sc.tlas2);
    const float4 *node = base + 4u * idx;
    float4 n0 = node[0], n1 = node[1], n2 = node[2], n3 = node[3];
    // Keep the four 16-byte loads of the node together ahead of the type
    // branch: left alone, the compiler loads the type word first and the
    // rest inside each branch, two dependent trips to memory per node.
    pin_node(n0, n1, n2, n3);
    const uint32_t type = as_u32(n2.w);
    if ((F & kHasSpheres) && type == 1u) {
        const uint32_t tid = as_u32(n3.w) * 2u;
        const float t = sphere_hit(xform_ray(sc.transforms + 4u * tid, tr.ray), t_min, tr.t_max);
        if (t < tr.t_max) {
            tr.hit = true;
            tr.t_max = t;
            tr.h.t = t;
            tr.h.material = as_u32(n1.w);
            tr.h.node_type = 1u;
            tr.h.inv_tid = tid;
        }
    } else if ((F & kHasMeshes) && type == 2u) {
        tr.in_tlas = false;
        if (tr.top < cap) stack[(tr.top++) * stride] = kFinishedTraverseBlas;
        tr.inv_tid = as_u32(n3.w) * 2u;
        tr.material_index = as_u32(n1.w);
        tr.ray = xform_ray(sc.transforms + 4u * tr.inv_tid, tr.ray);
        tr.inv = safe_invdir(tr.ray.d);
        tr.ox = (-tr.ray.o) * tr.inv;
        tr.addr = as_u32(n0.w);
        return true;
    } else if ((F & kHasMeshes) && type == 3u) {
        float u = 0.0f, v = 0.0f;
        const float t = triangle_hit(tr.ray, xyz(n0), xyz(n1), xyz(n2), t_min, tr.t_max, u, v);
        if (t < tr.t_max) {
            tr.hit = true;
            tr.t_max = t;
            tr.h.t = t;
            tr.h.material = tr.material_index;
            tr.h.node_type = 2u;
            tr.h.inv_tid = tr.inv_tid;
            tr.h.tri_id = as_u32(n0.w) * 3u;
            tr.h.u = u;
            tr.h.v = v;
        }
    }
    return trav_pop<F>(sc, world, stack, stride, tr);
}

template <uint32_t F, uint32_t W>
__device__ __forceinline__ bool trav_step(const DevScene &sc, float t_min, const Ray &world,
                                          uint32_t *stack, uint32_t stride, Traversal &tr)
{
    if (tr.addr & kLeafRef) return trav_leaf<F, W>(sc, t_min, world, stack, stride, tr);
    return trav_internal<F, W>(sc, t_min, world, stack, stride, tr);
}

// The whole of bvh_hit for one ray. With STEPS > 1 a lane visits up to STEPS
// internal nodes in a row and then, if it stands at one, a leaf: lanes of a
// wave that stand at leaves then test them together instead of each leaf
// visit splitting the wave. Every lane still meets its own nodes in the same
// order, so the result does not depend on STEPS.
template <uint32_t F, uint32_t W, uint32_t STEPS = 1>
__device__ __forceinline__ bool bvh_hit(const DevScene &sc, float t_min, const Ray &world,
                                        uint32_t *stack, uint32_t stride, Hit &h)
{
    Traversal tr;
    trav_begin(W == 4u ? sc.tlas4_root : sc.tlas2_root, world, tr);
    if constexpr (STEPS <= 1u) {
        while (trav_step<F, W>(sc, t_min, world, stack, stride, tr)) {
        }
    } else {
        bool on = true;
        while (on) {
            for (uint32_t k = 0; k < STEPS && on && !(tr.addr & kLeafRef); k++)
                on = trav_internal<F, W>(sc, t_min, world, stack, stride, tr);
            if (on && (tr.addr & kLeafRef)) on = trav_leaf<F, W>(sc, t_min, world, stack, stride, tr);
        }
    }
    h = tr.h;
    return tr.hit;
}

// bvh_hit for the 64 rays of a wave at once (BVH2). The wave walks the tree
// one node at a time on one stack (entry i held by lane i of three registers:
// the node and the 64-bit mask of the lanes that entered it). At an internal
// node each active lane tests both child boxes with its own ray and t_max;
// the wave goes on with the right child for the lanes whose ray hit it and
// pushes the left child for the lanes whose ray hit that. So a lane is active
// in exactly the nodes its own bvh_hit visits, meets them in the same order
// with the same t_max, and ends with bvh_hit's result bit for bit. For
// coherent rays (a camera tile) the wave visits few more nodes than one ray
// does, with every lane busy and every node load at one address. Every lane
// of the wave calls it (valid = false for a lane without a ray); `cap` is the
// scene's BVH2 stack bound (at most 64).
template <uint32_t F>
__device__ __forceinline__ bool bvh_hit_packet(const DevScene &sc, float t_min, const Ray &world, bool valid,
                                               uint32_t cap, Hit &h)
{
    Traversal tr;
    trav_begin(sc.tlas2_root, world, tr);
    const uint32_t lane = __lane_id();
    uint64_t active = lanes_where(valid);
    uint32_t addr = sc.tlas2_root;
    bool in_tlas = true;
    uint32_t top = 0;
    uint32_t s_node = 0, s_lo = 0, s_hi = 0;  // the wave's stack, entry i in lane i
    auto push = [&](uint32_t node, uint64_t mask) {
        if (top >= cap) return;
        const bool mine = lane == top;  // lane `top` holds the new entry
        s_node = mine ? node : s_node;
        s_lo = mine ? static_cast<uint32_t>(mask) : s_lo;
        s_hi = mine ? static_cast<uint32_t>(mask >> 32) : s_hi;
        top++;
    };
    while (active) {
        const bool me = (active >> lane) & 1ull;
        if (!(addr & kLeafRef)) {
            const float4 *node = (((F & kHasMeshes) && !in_tlas) ? sc.blas2 : sc.tlas2) + 4u * addr;
            const float4 n0 = node[0], n1 = node[1], n2 = node[2], n3 = node[3];
            const bool l = me && aabb_hit(make_float3(n0.x, n0.y, n0.z), make_float3(n1.x, n1.y, n1.z), tr.inv,
                                          tr.ox, t_min, tr.t_max);
            const bool r = me && aabb_hit(make_float3(n2.x, n2.y, n2.z), make_float3(n3.x, n3.y, n3.z), tr.inv,
                                          tr.ox, t_min, tr.t_max);
            const uint64_t ml = lanes_where(l), mr = lanes_where(r);
            const uint32_t left = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(as_u32(n0.w))));
            const uint32_t right = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(as_u32(n1.w))));
            if (mr) {
                if (ml) push(left, ml);
                addr = right;
                active = mr;
                continue;
            }
            if (ml) {
                addr = left;
                active = ml;
                continue;
            }
        } else {
            const uint32_t idx = addr & ~kLeafRef;
            const float4 *node = (((F & kHasMeshes) && !in_tlas) ? sc.blas2 : sc.tlas2) + 4u * idx;
            const float4 n0 = node[0], n1 = node[1], n2 = node[2], n3 = node[3];
            const uint32_t type = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(as_u32(n2.w))));
            if ((F & kHasSpheres) && type == 1u) {
                if (me) {
                    const uint32_t tid = as_u32(n3.w) * 2u;
                    const float t = sphere_hit(xform_ray(sc.transforms + 4u * tid, tr.ray), t_min, tr.t_max);
                    if (t < tr.t_max) {
                        tr.hit = true;
                        tr.t_max = t;
                        tr.h.t = t;
                        tr.h.material = as_u32(n1.w);
                        tr.h.node_type = 1u;
                        tr.h.inv_tid = tid;
                    }
                }
            } else if ((F & kHasMeshes) && type == 2u) {
                if (me) {
                    tr.inv_tid = as_u32(n3.w) * 2u;
                    tr.material_index = as_u32(n1.w);
                    tr.ray = xform_ray(sc.transforms + 4u * tr.inv_tid, tr.ray);
                    tr.inv = safe_invdir(tr.ray.d);
                    tr.ox = (-tr.ray.o) * tr.inv;
                }
                push(kFinishedTraverseBlas, 0ull);
                in_tlas = false;
                addr = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(as_u32(n0.w))));
                continue;  // the same lanes go on into the BLAS
            } else if ((F & kHasMeshes) && type == 3u) {
                if (me) {
                    float u = 0.0f, v = 0.0f;
                    const float t = triangle_hit(tr.ray, xyz(n0), xyz(n1), xyz(n2), t_min, tr.t_max, u, v);
                    if (t < tr.t_max) {
                        tr.hit = true;
                        tr.t_max = t;
                        tr.h.t = t;
                        tr.h.material = tr.material_index;
                        tr.h.node_type = 2u;
                        tr.h.inv_tid = tr.inv_tid;
                        tr.h.tri_id = as_u32(n0.w) * 3u;
                        tr.h.u = u;
                        tr.h.v = v;
                    }
                }
            }
        }
        // next pending entry; leaving a BLAS restores every lane's world ray
        // (a no-op for the lanes that did not enter it)
        active = 0ull;
        while (top) {
            --top;
            const uint32_t a = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(s_node),
                                                                               static_cast<int>(top)));
            if ((F & kHasMeshes) && a == kFinishedTraverseBlas) {
                in_tlas = true;
                tr.ray = world;
                tr.inv = tr.winv;
                tr.ox = tr.wox;
                continue;
            }
            addr = a;
            active = (uint64_t(static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(s_hi),
                                                                               static_cast<int>(top))))
                      << 32) |
                     static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(s_lo), static_cast<int>(top)));
            break;
        }
    }
    h = tr.h;
    return tr.hit;
}


// ---------------------------------------------------------------------------
// Shading: material.wgsl, hitrecord.wgsl, pathtracer.wgsl:95-161
// ---------------------------------------------------------------------------

struct Material {
    f3 albedo;
    uint32_t texture;
    float fuzz;
    float ior;
    uint32_t type;
};

__device__ __forceinline__ Material load_material(const DevScene &sc, uint32_t i)
{
    const uint4 a = sc.materials[2u * i], b = sc.materials[2u * i + 1u];
    Material m;
    m.albedo = f3{as_f32(a.x), as_f32(a.y), as_f32(a.z)};
    m.texture = a.w;
    m.fuzz = as_f32(b.x);
    m.ior = as_f32(b.y);
    m.type = b.z;
    return m;
}

__device__ __forceinline__ float half_to_float(uint32_t h)
{
    const uint32_t sign = (h >> 15) << 31;
    const uint32_t e = (h >> 10) & 0x1fu, man = h & 0x3ffu;
    if (e == 0u) {
        const float f = static_cast<float>(man) * 5.9604644775390625e-08f;
        return sign ? -f : f;
    }
    if (e == 31u) return as_f32(sign | 0x7f800000u | (man << 13));
    return as_f32(sign | ((e + 112u) << 23) | (man << 13));
}

// textureSampleLevel with the default WebGPU sampler: nearest, clamp to edge,
// level 0 (webgpu.zig:812-825); rgba8unorm reads as byte / 255.
__device__ __forceinline__ f3 sample_texture(const DevScene &sc, uint32_t id, float u, float v)
{
    const DevTexture tx = sc.textures[id];
    float fx = floorf(u * static_cast<float>(tx.width));
    float fy = floorf(v * static_cast<float>(tx.height));
    fx = fminf(fmaxf(fx, 0.0f), static_cast<float>(tx.width - 1u));
    fy = fminf(fmaxf(fy, 0.0f), static_cast<float>(tx.height - 1u));
    const uint64_t texel = static_cast<uint64_t>(static_cast<uint32_t>(fy)) * tx.width +
                           static_cast<uint32_t>(fx);
    float c[3] = {0.0f, 0.0f, 0.0f};
    if (tx.hdr) {
        const uint16_t *d =
            reinterpret_cast<const uint16_t *>(sc.texels + tx.offset) + texel * tx.components;
        if (tx.components == 4u) {
            const uint2 q = *reinterpret_cast<const uint2 *>(d);
            c[0] = half_to_float(q.x & 0xffffu);
            c[1] = half_to_float(q.x >> 16);
            c[2] = half_to_float(q.y & 0xffffu);
        } else {
            for (uint32_t k = 0; k < tx.components; k++) c[k] = half_to_float(d[k]);
        }
    } else {
        const uint8_t *d = sc.texels + tx.offset + texel * tx.components;
        if (tx.components == 4u) {
            const uint32_t q = *reinterpret_cast<const uint32_t *>(d);
            c[0] = static_cast<float>(q & 0xffu) / 255.0f;
            c[1] = static_cast<float>((q >> 8) & 0xffu) / 255.0f;
            c[2] = static_cast<float>((q >> 16) & 0xffu) / 255.0f;
        } else {
            for (uint32_t k = 0; k < tx.components; k++) c[k] = static_cast<float>(d[k]) / 255.0f;
        }
    }
    return f3{c[0], c[1], c[2]};
}

__device__ __forceinline__ bool near_zero(f3 e)
{
    return fabsf(e.x) < 1e-8f && fabsf(e.y) < 1e-8f && fabsf(e.z) < 1e-8f;
}

// WGSL reflect / refract
__device__ __forceinline__ f3 reflect_dir(f3 v, f3 n) { return v - (2.0f * dot(n, v)) * n; }

__device__ __forceinline__ f3 refract_dir(f3 e1, f3 e2, float eta)
{
    const float d = dot(e2, e1);
    const float k = 1.0f - eta * eta * (1.0f - d * d);
    if (k < 0.0f) return f3{0.0f, 0.0f, 0.0f};
    return eta * e1 - (eta * d + sqrtf(k)) * e2;
}

// The shading point of a closest hit (pathtracer.wgsl:114-135,
// hitrecord.wgsl:10-18): position, front-facing shading normal, material and
// texture coordinates (uv is only computed when a texture will be read).
struct Surface {
    f3 p;
    f3 normal;
    bool front;
    bool textured;
    float tu, tv;
    Material m;
};

// The same operations as the first half of shade() (which keeps its own
// copy: sharing this function changed the hot kernel's register allocation).
template <uint32_t F>
__device__ __forceinline__ Surface surface(const DevScene &sc, uint32_t textures_count, uint32_t tex_base,
                                           const Ray &ray, const Hit &h)
{
    Surface s;
    s.p = ray.o + h.t * ray.d;
    s.m = load_material(sc, h.material);
    const Material &m = s.m;
    s.textured =
        (F & kHasTextures) && m.texture - tex_base < textures_count && m.texture < sc.textures_len;
    s.tu = 0.0f;
    s.tv = 0.0f;
    f3 outward;
    if ((F & kHasSpheres) && (!(F & kHasMeshes) || h.node_type == 1u)) {
        // sphere: outward = normalize(p - M * (0,0,0,1))
        const float4 c3 = sc.transforms[4u * (h.inv_tid + 1u) + 3u];
        const float4 c0 = sc.transforms[4u * (h.inv_tid + 1u)];
        const float4 c1 = sc.transforms[4u * (h.inv_tid + 1u) + 1u];
        const float4 c2 = sc.transforms[4u * (h.inv_tid + 1u) + 2u];
        const f3 center = f3{c0.x * 0.0f + c1.x * 0.0f + c2.x * 0.0f + c3.x * 1.0f,
                             c0.y * 0.0f + c1.y * 0.0f + c2.y * 0.0f + c3.y * 1.0f,
                             c0.z * 0.0f + c1.z * 0.0f + c2.z * 0.0f + c3.z * 1.0f};
        outward = normalize(s.p - center);
        if (s.textured) {  // uv only reaches the result through a texture fetch
            const float theta = acos_f(-outward.y);
            const float phi = atan2_f(-outward.z, outward.x) + kPi;
            s.tu = phi / (2.0f * kPi);
            s.tv = theta / kPi;
        }
    } else {
        // mesh: barycentric shading normal and uv, normal by transpose(inverse)
        const uint32_t i0 = sc.normal_indices[h.tri_id], i1 = sc.normal_indices[h.tri_id + 1u],
                       i2 = sc.normal_indices[h.tri_id + 2u];
        const f3 n0 = xyz(sc.normals[i0]), n1 = xyz(sc.normals[i1]), n2 = xyz(sc.normals[i2]);
        const float w = 1.0f - h.u - h.v;
        const f3 nrm = (w * n0 + h.u * n1) + h.v * n2;
        if (s.textured) {
            const float2 uv0 = sc.uvs[sc.uv_indices[h.tri_id]];
            const float2 uv1 = sc.uvs[sc.uv_indices[h.tri_id + 1u]];
            const float2 uv2 = sc.uvs[sc.uv_indices[h.tri_id + 2u]];
            s.tu = w * uv0.x + h.u * uv1.x + h.v * uv2.x;
            s.tv = w * uv0.y + h.u * uv1.y + h.v * uv2.y;
        }
        outward = normalize(xform_normal(sc.transforms + 4u * h.inv_tid, nrm));
    }
    // hitrecord.wgsl:10-18
    if (dot(ray.d, outward) > 0.0f) {
        s.normal = -outward;
        s.front = false;
    } else {
        s.normal = outward;
        s.front = true;
    }
    return s;
}

// Shading of one closest hit: fills the next ray and the attenuation, or the
// emitted radiance when the path ends (pathtracer.wgsl:114-157).
// Returns true when the path continues.
// Texture lookup (material.wgsl:91-99): a material's texture is sampled when
// its index, counted from `tex_base`, is below `textures_count`. tex_base is
// the scene's first texture in the shared texture table (non-zero when
// several scenes share one device scene; their materials then hold indices
// into the shared table, and 0xffffffff for "none"). Textures beyond what was
// uploaded read as "no texture" (the reference would index its binding array
// out of range there).
template <uint32_t F>
__device__ __forceinline__ bool shade(const DevScene &sc, uint32_t textures_count, uint32_t tex_base,
                                      const Ray &ray, const Hit &h, uint32_t &rng, Ray &next, f3 &weight)
{
    const f3 p = ray.o + h.t * ray.d;
    const Material m = load_material(sc, h.material);
    const bool textured =
        (F & kHasTextures) && m.texture - tex_base < textures_count && m.texture < sc.textures_len;
    f3 normal;
    bool front;
    float tu = 0.0f, tv = 0.0f;
    f3 outward;
    if ((F & kHasSpheres) && (!(F & kHasMeshes) || h.node_type == 1u)) {
        // sphere: outward = normalize(p - M * (0,0,0,1))
        const float4 c3 = sc.transforms[4u * (h.inv_tid + 1u) + 3u];
        const float4 c0 = sc.transforms[4u * (h.inv_tid + 1u)];
        const float4 c1 = sc.transforms[4u * (h.inv_tid + 1u) + 1u];
        const float4 c2 = sc.transforms[4u * (h.inv_tid + 1u) + 2u];
        const f3 center = f3{c0.x * 0.0f + c1.x * 0.0f + c2.x * 0.0f + c3.x * 1.0f,
                             c0.y * 0.0f + c1.y * 0.0f + c2.y * 0.0f + c3.y * 1.0f,
                             c0.z * 0.0f + c1.z * 0.0f + c2.z * 0.0f + c3.z * 1.0f};
        outward = normalize(p - center);
        if (textured) {  // uv only reaches the result through a texture fetch
            const float theta = acos_f(-outward.y);
            const float phi = atan2_f(-outward.z, outward.x) + kPi;
            tu = phi / (2.0f * kPi);
            tv = theta / kPi;
        }
    } else {
        // mesh: barycentric shading normal and uv, normal by transpose(inverse)
        const uint32_t i0 = sc.normal_indices[h.tri_id], i1 = sc.normal_indices[h.tri_id + 1u],
                       i2 = sc.normal_indices[h.tri_id + 2u];
        const f3 n0 = xyz(sc.normals[i0]), n1 = xyz(sc.normals[i1]), n2 = xyz(sc.normals[i2]);
        const float w = 1.0f - h.u - h.v;
        const f3 nrm = (w * n0 + h.u * n1) + h.v * n2;
        if (textured) {
            const float2 uv0 = sc.uvs[sc.uv_indices[h.tri_id]];
            const float2 uv1 = sc.uvs[sc.uv_indices[h.tri_id + 1u]];
            const float2 uv2 = sc.uvs[sc.uv_indices[h.tri_id + 2u]];
            tu = w * uv0.x + h.u * uv1.x + h.v * uv2.x;
            tv = w * uv0.y + h.u * uv1.y + h.v * uv2.y;
        }
        outward = normalize(xform_normal(sc.transforms + 4u * h.inv_tid, nrm));
    }
    // hitrecord.wgsl:10-18
    if (dot(ray.d, outward) > 0.0f) {
        normal = -outward;
        front = false;
    } else {
        normal = outward;
        front = true;
    }
    if (m.type <= 1u) {
        // Lambertian (material.wgsl:9-20) and metal (:22-27) draw the same
        // three numbers for random_in_unit_sphere; one code path serves both
        // kinds of lanes in a wave.
        const f3 s = random_in_unit_sphere(rng);
        f3 dir;
        bool cont;
        if (m.type == 0u) {
            dir = normal + normalize(s);
            if (near_zero(dir)) dir = normal;
            cont = true;
        } else {
            dir = reflect_dir(normalize(ray.d), normal) + m.fuzz * s;
            cont = dot(dir, normal) > 0.0f;  // absorbed below the surface
        }
        next.o = p;
        next.d = dir;
        if (cont) {
            weight = textured ? sample_texture(sc, m.texture, tu, tv) : m.albedo;
            return true;
        }
        weight = f3{0.0f, 0.0f, 0.0f};  // material_emit of a metal
        return false;
    }
    switch (m.type) {
    case 2u: {  // material.wgsl:36-59
        const float ratio = front ? (1.0f / m.ior) : m.ior;
        const f3 ud = normalize(ray.d);
        const float cos_theta = fminf(dot(-ud, normal), 1.0f);
        const float sin_theta = sqrtf(1.0f - cos_theta * cos_theta);
        const bool cannot_refract = ratio * sin_theta > 1.0f;
        bool reflect = cannot_refract;
        if (!reflect) {  // the draw happens only when refraction is possible
            float r0 = (1.0f - ratio) / (1.0f + ratio);
            r0 = r0 * r0;
            const float refl = r0 + (1.0f - r0) * pow_f(1.0f - cos_theta, 5.0f);
            reflect = refl > rng_f32(rng);
        }
        next.o = p;
        next.d = reflect ? reflect_dir(ud, normal) : refract_dir(ud, normal, ratio);
        weight = f3{1.0f, 1.0f, 1.0f};
        return true;
    }
    default:  // diffuse light (3) emits; anything else emits black
        weight = (m.type == 3u) ? (textured ? sample_texture(sc, m.texture, tu, tv) : m.albedo)
                                : f3{0.0f, 0.0f, 0.0f};
        return false;
    }
}


// The material part of shade() (material.wgsl:9-89) at a shading point from
// surface(): the same operations, for callers that need the shading point's
// data before scattering.
template <uint32_t F>
__device__ __forceinline__ bool scatter(const DevScene &sc, const Surface &srf, const Ray &ray, uint32_t &rng,
                                        Ray &next, f3 &weight)
{
    const Material &m = srf.m;
    const f3 &p = srf.p;
    const f3 &normal = srf.normal;
    const bool front = srf.front;
    const bool textured = srf.textured;
    const float tu = srf.tu, tv = srf.tv;
    if (m.type <= 1u) {
        // Lambertian (material.wgsl:9-20) and metal (:22-27) draw the same
        // three numbers for random_in_unit_sphere; one code path serves both
        // kinds of lanes in a wave.
        const f3 s = random_in_unit_sphere(rng);
        f3 dir;
        bool cont;
        if (m.type == 0u) {
            dir = normal + normalize(s);
            if (near_zero(dir)) dir = normal;
            cont = true;
        } else {
            dir = reflect_dir(normalize(ray.d), normal) + m.fuzz * s;
            cont = dot(dir, normal) > 0.0f;  // absorbed below the surface
        }
        next.o = p;
        next.d = dir;
        if (cont) {
            weight = textured ? sample_texture(sc, m.texture, tu, tv) : m.albedo;
            return true;
        }
        weight = f3{0.0f, 0.0f, 0.0f};  // material_emit of a metal
        return false;
    }
    switch (m.type) {
    case 2u: {  // material.wgsl:36-59
        const float ratio = front ? (1.0f / m.ior) : m.ior;
        const f3 ud = normalize(ray.d);
        const float cos_theta = fminf(dot(-ud, normal), 1.0f);
        const float sin_theta = sqrtf(1.0f - cos_theta * cos_theta);
        const bool cannot_refract = ratio * sin_theta > 1.0f;
        bool reflect = cannot_refract;
        if (!reflect) {  // the draw happens only when refraction is possible
            float r0 = (1.0f - ratio) / (1.0f + ratio);
            r0 = r0 * r0;
            const float refl = r0 + (1.0f - r0) * pow_f(1.0f - cos_theta, 5.0f);
            reflect = refl > rng_f32(rng);
        }
        next.o = p;
        next.d = reflect ? reflect_dir(ud, normal) : refract_dir(ud, normal, ratio);
        weight = f3{1.0f, 1.0f, 1.0f};
        return true;
    }
    default:  // diffuse light (3) emits; anything else emits black
        weight = (m.type == 3u) ? (textured ? sample_texture(sc, m.texture, tu, tv) : m.albedo)
                                : f3{0.0f, 0.0f, 0.0f};
        return false;
    }
}

// Two floats to packed bf16 (first in the low half) with gfx950's
// v_cvt_pk_bf16_f32, which rounds to nearest even. For every non-NaN input
// that is the software rounding (u + 0x7fff + ((u >> 16) & 1)) >> 16 of the
// radiance cache's numerics; a NaN input may come out with another NaN
// payload. orn_eval_math op 5 evaluates it for the tests.
__device__ __forceinline__ uint32_t pack_bf16_rne(float a, float b)
{
    typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
    const bf16x2_t v = {static_cast<__bf16>(a), static_cast<__bf16>(b)};
    return __builtin_bit_cast(uint32_t, v);
}

// Low 32 bits of the device's constant 100 MHz clock (wave-uniform).
__device__ __forceinline__ uint32_t device_ticks() { return static_cast<uint32_t>(__builtin_amdgcn_s_memrealtime()); }

// Quarter-octave class of a cost (queues ordered by cost): 4 * floor(log2(c +
// 1)) plus the two bits below the leading one of c + 1; non-decreasing in c,
// below 128.
constexpr uint32_t kCostClasses = 128;
__device__ __forceinline__ uint32_t cost_class(uint32_t c)
{
    const uint32_t v = c == 0xffffffffu ? c : c + 1u;
    const uint32_t e = 31u - static_cast<uint32_t>(__clz(v));
    const uint32_t f = e >= 2u ? (v >> (e - 2u)) & 3u : (v << (2u - e)) & 3u;
    return 4u * e + f;
}

}  // namespace orn
