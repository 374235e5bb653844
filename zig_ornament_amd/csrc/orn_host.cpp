// orn_host.cpp -- the ornament host layer in C++ (include/ornament.h).
//
// The reference keeps this layer in Zig: scene factories (src/scene.zig),
// camera (src/camera.zig), render state (src/state.zig), the TLAS/BLAS build
// (src/bvh.zig) and the backend driver loop (wgpu_backend/path_tracer.zig).
// Here it is restated in C++ (no Zig toolchain is assumed) with
// the same names, argument meaning and error behaviour, and drives the gfx950
// kernels through include/ornament_hip.h.
//
// Host arithmetic follows zmath's row-vector conventions in f32 with one
// rounding per written operation (compiled with -ffp-contract=off), except
// matrix inversion, which is done in double and rounded once to f32.
#include <array>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <memory>
#include <mutex>
#include <new>
#include <unordered_set>
#include <vector>

#include "../../include/ornament.h"

namespace orn_host {

using Vec = std::array<float, 4>;   // zmath.Vec
using Mat = std::array<float, 16>;  // zmath.Mat, rows contiguous

constexpr float kPi = 3.1415927410125732f;

Vec vec(float x, float y, float z, float w) { return Vec{x, y, z, w}; }
Vec load_vec(const float *p) { return Vec{p[0], p[1], p[2], p[3]}; }
Vec sub(const Vec &a, const Vec &b) { return Vec{a[0] - b[0], a[1] - b[1], a[2] - b[2], a[3] - b[3]}; }
Vec scale(float s, const Vec &a) { return Vec{s * a[0], s * a[1], s * a[2], s * a[3]}; }
float dot3(const Vec &a, const Vec &b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
// zmath cross3: w is masked to 0
Vec cross3(const Vec &a, const Vec &b)
{
    return Vec{a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0], 0.0f};
}
// zmath normalize3: v * 1 / sqrt(dot3(v, v)), evaluated as v / sqrt(...)
Vec normalize3(const Vec &v)
{
    const float len = std::sqrt(dot3(v, v));
    return Vec{v[0] / len, v[1] / len, v[2] / len, v[3] / len};
}
// zmath normalize4 with dot4 = (z*z + w*w) + (x*x + y*y)
Vec normalize4(const Vec &v)
{
    const float len = std::sqrt((v[2] * v[2] + v[3] * v[3]) + (v[0] * v[0] + v[1] * v[1]));
    return Vec{v[0] / len, v[1] / len, v[2] / len, v[3] / len};
}
Vec vmin(const Vec &a, const Vec &b)
{
    Vec r;
    for (int i = 0; i < 4; i++) r[i] = std::fmin(a[i], b[i]);
    return r;
}
Vec vmax(const Vec &a, const Vec &b)
{
    Vec r;
    for (int i = 0; i < 4; i++) r[i] = std::fmax(a[i], b[i]);
    return r;
}

Mat identity()
{
    Mat m{};
    m[0] = m[5] = m[10] = m[15] = 1.0f;
    return m;
}
Mat scaling(float x, float y, float z)
{
    Mat m = identity();
    m[0] = x;
    m[5] = y;
    m[10] = z;
    return m;
}
Mat translation(float x, float y, float z)
{
    Mat m = identity();
    m[12] = x;
    m[13] = y;
    m[14] = z;
    return m;
}
// zmath mul(Mat, Mat): row_i = (x*b0 + z*b2) + (y*b1 + w*b3)
Mat mul(const Mat &a, const Mat &b)
{
    Mat r;
    for (int i = 0; i < 4; i++) {
        const float x = a[4 * i], y = a[4 * i + 1], z = a[4 * i + 2], w = a[4 * i + 3];
        for (int j = 0; j < 4; j++) r[4 * i + j] = (x * b[j] + z * b[8 + j]) + (y * b[4 + j] + w * b[12 + j]);
    }
    return r;
}
// zmath mul(Vec, Mat): ((x*m0 + y*m1) + z*m2) + w*m3
Vec mul(const Vec &v, const Mat &m)
{
    Vec r;
    for (int j = 0; j < 4; j++) r[j] = ((v[0] * m[j] + v[1] * m[4 + j]) + v[2] * m[8 + j]) + v[3] * m[12 + j];
    return r;
}
// zmath matFromQuat (rotation of row vectors by a unit quaternion)
Mat mat_from_quat(const Vec &q)
{
    const float x = q[0], y = q[1], z = q[2], w = q[3];
    const float x2 = x + x, y2 = y + y, z2 = z + z;
    const float xx = x * x2, yy = y * y2, zz = z * z2;
    const float xy = x * y2, xz = x * z2, yz = y * z2;
    const float wx = w * x2, wy = w * y2, wz = w * z2;
    return Mat{(1.0f - yy) - zz, xy + wz, xz - wy, 0.0f,
               xy - wz, (1.0f - xx) - zz, yz + wx, 0.0f,
               xz + wy, yz - wx, (1.0f - xx) - yy, 0.0f,
               0.0f, 0.0f, 0.0f, 1.0f};
}
// Inverse through the adjugate built from 2x2 minors, in double, rounded
// once to f32; zeros are written as +0. zmath.inverse's own f32 bit pattern
// is not reproducible without zmath; this is the closest-to-exact choice
// (exact zeros stay exact).
bool inverse(const Mat &m, Mat &out)
{
    double a[16];
    for (int i = 0; i < 16; i++) a[i] = m[i];
    auto A = [&](int r, int c) { return a[4 * r + c]; };
    const double s0 = A(0, 0) * A(1, 1) - A(1, 0) * A(0, 1);
    const double s1 = A(0, 0) * A(1, 2) - A(1, 0) * A(0, 2);
    const double s2 = A(0, 0) * A(1, 3) - A(1, 0) * A(0, 3);
    const double s3 = A(0, 1) * A(1, 2) - A(1, 1) * A(0, 2);
    const double s4 = A(0, 1) * A(1, 3) - A(1, 1) * A(0, 3);
    const double s5 = A(0, 2) * A(1, 3) - A(1, 2) * A(0, 3);
    const double c5 = A(2, 2) * A(3, 3) - A(3, 2) * A(2, 3);
    const double c4 = A(2, 1) * A(3, 3) - A(3, 1) * A(2, 3);
    const double c3 = A(2, 1) * A(3, 2) - A(3, 1) * A(2, 2);
    const double c2 = A(2, 0) * A(3, 3) - A(3, 0) * A(2, 3);
    const double c1 = A(2, 0) * A(3, 2) - A(3, 0) * A(2, 2);
    const double c0 = A(2, 0) * A(3, 1) - A(3, 0) * A(2, 1);
    const double det = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
    if (det == 0.0) return false;
    const double adj[16] = {
        A(1, 1) * c5 - A(1, 2) * c4 + A(1, 3) * c3,  -A(0, 1) * c5 + A(0, 2) * c4 - A(0, 3) * c3,
        A(3, 1) * s5 - A(3, 2) * s4 + A(3, 3) * s3,  -A(2, 1) * s5 + A(2, 2) * s4 - A(2, 3) * s3,
        -A(1, 0) * c5 + A(1, 2) * c2 - A(1, 3) * c1, A(0, 0) * c5 - A(0, 2) * c2 + A(0, 3) * c1,
        -A(3, 0) * s5 + A(3, 2) * s2 - A(3, 3) * s1, A(2, 0) * s5 - A(2, 2) * s2 + A(2, 3) * s1,
        A(1, 0) * c4 - A(1, 1) * c2 + A(1, 3) * c0,  -A(0, 0) * c4 + A(0, 1) * c2 - A(0, 3) * c0,
        A(3, 0) * s4 - A(3, 1) * s2 + A(3, 3) * s0,  -A(2, 0) * s4 + A(2, 1) * s2 - A(2, 3) * s0,
        -A(1, 0) * c3 + A(1, 1) * c1 - A(1, 2) * c0, A(0, 0) * c3 - A(0, 1) * c1 + A(0, 2) * c0,
        -A(3, 0) * s3 + A(3, 1) * s1 - A(3, 2) * s0, A(2, 0) * s3 - A(2, 1) * s1 + A(2, 2) * s0,
    };
    for (int i = 0; i < 16; i++) out[i] = static_cast<float>(adj[i] / det) + 0.0f;
    return true;
}

// math.zig:13-41
Mat rotation_between_vectors(const Vec &a, const Vec &b)
{
    auto approx = [](float x, float y) { return std::fabs(x - y) <= 0.0000001f; };
    const float k_cos_theta = dot3(a, b);
    if (approx(k_cos_theta, 1.0f)) return identity();
    const float k = std::sqrt(dot3(a, a) * dot3(b, b));
    if (approx(k_cos_theta / k, -1.0f)) {
        Vec o = cross3(a, vec(1.0f, 0.0f, 0.0f, 0.0f));
        if (approx(dot3(o, o), 0.0f)) o = cross3(a, vec(0.0f, 1.0f, 0.0f, 0.0f));
        o = normalize3(o);
        return mat_from_quat(vec(o[0], o[1], o[2], 0.0f));
    }
    const Vec v = cross3(a, b);
    return mat_from_quat(normalize4(vec(v[0], v[1], v[2], k + k_cos_theta)));
}

// math.zig:43-47
std::array<float, 2> sphere_tex_coord(const Vec &p)
{
    const float theta = static_cast<float>(std::acos(static_cast<double>(-p[1])));
    const float phi = static_cast<float>(std::atan2(static_cast<double>(-p[2]), static_cast<double>(p[0]))) + kPi;
    return {phi / (2.0f * kPi), theta / kPi};
}

struct Aabb {
    Vec min, max;
    void grow(const Vec &p)
    {
        min = vmin(min, p);
        max = vmax(max, p);
    }
    // aabb.zig:17-45: the eight corners through the matrix
    static Aabb transform(const Mat &m, const Aabb &b)
    {
        const Vec p[8] = {b.min,
                          vec(b.max[0], b.min[1], b.min[2], 1.0f),
                          vec(b.min[0], b.max[1], b.min[2], 1.0f),
                          vec(b.min[0], b.min[1], b.max[2], 1.0f),
                          vec(b.min[0], b.max[1], b.max[2], 1.0f),
                          vec(b.max[0], b.max[1], b.min[2], 1.0f),
                          vec(b.max[0], b.min[1], b.max[2], 1.0f),
                          b.max};
        Aabb r{mul(p[0], m), mul(p[0], m)};
        for (int i = 1; i < 8; i++) r.grow(mul(p[i], m));
        return r;
    }
};

}  // namespace orn_host

using namespace orn_host;

// ---------------------------------------------------------------------------
// Scene objects (material.zig, texture.zig, sphere.zig, mesh.zig,
// mesh_instance.zig), owned by their scene.
// ---------------------------------------------------------------------------
struct ornament_texture {
    std::vector<uint8_t> data;
    uint32_t width, height, num_components, bytes_per_component, bytes_per_row;
    bool is_hdr;
    float gamma;
    int64_t texture_id = -1;  // ?u32, set by the BVH build
};

struct ornament_material {
    Vec albedo{};                         // Color.vec
    ornament_texture *texture = nullptr;  // Color.texture
    float fuzz = 0.0f;                    // `undefined` in the reference for some types
    float ior = 0.0f;
    uint32_t type = 0;
    int64_t material_index = -1;  // ?u32, set by the BVH build
};

struct ornament_sphere {
    ornament_material *material;
    Mat transform;
    Aabb aabb;
};

struct ornament_mesh {
    std::vector<Vec> vertices;
    std::vector<uint32_t> vertex_indices;
    std::vector<Vec> normals;
    std::vector<uint32_t> normal_indices;
    std::vector<std::array<float, 2>> uvs;
    std::vector<uint32_t> uv_indices;
    Mat transform;
    ornament_material *material;
    int64_t bvh_id = -1;
    Aabb aabb, not_transformed_aabb;
};

struct ornament_mesh_instance {
    const ornament_mesh *mesh;
    ornament_material *material;
    Mat transform;
    Aabb aabb;
};

// camera.zig:4-114
struct Camera {
    Vec origin, lower_left_corner, horizontal, vertical, u, v, w;
    float lens_radius;
    float vfov, focus_dist, aspect_ratio;
    Vec lookfrom, lookat, vup;
    bool dirty;

    static Camera init(const Vec &lookfrom, const Vec &lookat, const Vec &vup, float aspect_ratio,
                       float vfov, float aperture, float focus_dist)
    {
        Camera c;
        const float theta = vfov * kPi / 180.0f;  // std.math.degreesToRadians
        const float h = static_cast<float>(std::tan(static_cast<double>(theta / 2.0f)));
        const float viewport_height = 2.0f * h;
        const float viewport_width = aspect_ratio * viewport_height;
        c.w = normalize3(sub(lookfrom, lookat));
        c.u = normalize3(cross3(vup, c.w));
        c.v = cross3(c.w, c.u);
        c.origin = lookfrom;
        c.horizontal = scale(focus_dist * viewport_width, c.u);
        c.vertical = scale(focus_dist * viewport_height, c.v);
        Vec half_h, half_v;
        for (int i = 0; i < 4; i++) {
            half_h[i] = c.horizontal[i] / 2.0f;
            half_v[i] = c.vertical[i] / 2.0f;
        }
        c.lower_left_corner = sub(sub(sub(c.origin, half_h), half_v), scale(focus_dist, c.w));
        c.lens_radius = aperture / 2.0f;
        c.focus_dist = focus_dist;
        c.vfov = vfov;
        c.aspect_ratio = aspect_ratio;
        c.lookfrom = lookfrom;
        c.lookat = lookat;
        c.vup = vup;
        c.dirty = true;
        return c;
    }
};

struct ornament_scene {
    Camera camera;
    std::vector<std::unique_ptr<ornament_sphere>> spheres;
    std::vector<std::unique_ptr<ornament_mesh>> meshes;
    std::vector<std::unique_ptr<ornament_mesh_instance>> mesh_instances;
    std::vector<std::unique_ptr<ornament_material>> materials;
    std::vector<std::unique_ptr<ornament_texture>> textures;
    std::vector<ornament_sphere *> attached_spheres;
    std::vector<ornament_mesh *> attached_meshes;
    std::vector<ornament_mesh_instance *> attached_mesh_instances;
};

// ---------------------------------------------------------------------------
// BVH build (bvh.zig)
// ---------------------------------------------------------------------------
namespace {

// std.rand.DefaultPrng of Zig 0.11: Xoshiro256 (xoshiro256++) seeded by
// SplitMix64. The reference shares one generator seeded with 1244 across all
// builds in a process (bvh.zig:12-13). By default this layer reseeds it per
// Bvh.init, so that a build does not depend on what was built before it in
// the same process; ornament_bvh_shared_generator(1) switches to one shared
// stream, as in the reference.
struct Xoshiro256 {
    uint64_t s[4];
    explicit Xoshiro256(uint64_t seed)
    {
        uint64_t g = seed;
        for (auto &w : s) {
            g += 0x9e3779b97f4a7c15ull;
            uint64_t z = g;
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
            w = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        const uint64_t r = rotl(s[0] + s[3], 23) + s[0];
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    // Random.intRangeAtMost(usize, 0, at_most) -> uintLessThan(at_most + 1):
    // Lemire's multiply-shift with the rejection threshold computed lazily
    uint64_t at_most(uint64_t m)
    {
        const uint64_t bound = m + 1u;
        unsigned __int128 prod = static_cast<unsigned __int128>(next()) * bound;
        uint64_t low = static_cast<uint64_t>(prod);
        if (low < bound) {
            uint64_t threshold = 0u - bound;
            if (threshold >= bound) {
                threshold -= bound;
                if (threshold >= bound) threshold %= bound;
            }
            while (low < threshold) {
                prod = static_cast<unsigned __int128>(next()) * bound;
                low = static_cast<uint64_t>(prod);
            }
        }
        return static_cast<uint64_t>(prod >> 64);
    }
};

// std.sort.heap of Zig 0.11 (unstable; the exact permutation decides node order)
template <typename T, typename Less>
void zig_heap_sort(T *items, size_t n, Less less)
{
    auto sift = [&](size_t root, size_t end) {
        size_t node = root;
        for (;;) {
            size_t child = 2 * node + 1;
            if (child >= end) break;
            if (child + 1 < end && less(items[child], items[child + 1])) child += 1;
            if (!less(items[node], items[child])) break;
            std::swap(items[node], items[child]);
            node = child;
        }
    };
    for (size_t i = n / 2; i > 0; i--) sift(i - 1, n);
    for (size_t i = n; i > 0; i--) {
        std::swap(items[0], items[i - 1]);
        sift(0, i - 1);
    }
}

struct Triangle {
    Vec v0, v1, v2;
    uint32_t triangle_index;
    Aabb aabb;
};

struct Leaf {
    enum Kind { Sphere, Mesh, MeshInstance } kind;
    void *ptr;
    const Aabb &aabb() const
    {
        switch (kind) {
        case Sphere: return static_cast<ornament_sphere *>(ptr)->aabb;
        case Mesh: return static_cast<ornament_mesh *>(ptr)->aabb;
        default: return static_cast<ornament_mesh_instance *>(ptr)->aabb;
        }
    }
};

orn_bvh_node internal_node(const Aabb &l, uint32_t left_id, const Aabb &r, uint32_t right_id)
{
    orn_bvh_node n{};
    std::memcpy(n.lmin_or_v0, l.min.data(), 12);
    n.left_or_custom_id = left_id;
    std::memcpy(n.lmax_or_v1, l.max.data(), 12);
    n.right_or_material_index = right_id;
    std::memcpy(n.rmin_or_v2, r.min.data(), 12);
    n.node_type = 0u;
    std::memcpy(n.rmax_or_v3, r.max.data(), 12);
    return n;
}

template <typename It, typename BoxOf>
Aabb bounding_box(It begin, It end, BoxOf box)
{
    const float inf = INFINITY;
    Aabb b{vec(inf, inf, inf, 1.0f), vec(-inf, -inf, -inf, 1.0f)};
    for (It i = begin; i != end; ++i) {
        b.min = vmin(b.min, box(*i).min);
        b.max = vmax(b.max, box(*i).max);
    }
    return b;
}

// The process's build generator when builds share one stream (see Xoshiro256).
struct SharedBvhPrng {
    std::mutex mutex;
    bool shared = false;
    uint64_t state[4] = {0, 0, 0, 0};
};
SharedBvhPrng g_bvh_prng;

}  // namespace

struct ornament_bvh {
    std::vector<orn_bvh_node> tlas_nodes, blas_nodes;
    std::vector<std::array<float, 4>> normals;
    std::vector<uint32_t> normal_indices;
    std::vector<std::array<float, 2>> uvs;
    std::vector<uint32_t> uv_indices;
    std::vector<Mat> transforms;
    std::vector<orn_material> materials;
    std::vector<ornament_texture *> textures;
    std::vector<orn_texture_desc> texture_descs;
    std::vector<uint32_t> texture_sources;
    bool row_major_transforms = false;

    Xoshiro256 rand{1244u};

    // bvh.zig:364-367
    void append_transform(const Mat &t)
    {
        Mat m = t;
        if (row_major_transforms)
            for (int i = 0; i < 4; i++)
                for (int j = i + 1; j < 4; j++) std::swap(m[4 * i + j], m[4 * j + i]);
        transforms.push_back(m);
    }

    // bvh.zig:369-383
    uint32_t material_index(ornament_material *m)
    {
        if (m->material_index >= 0) return static_cast<uint32_t>(m->material_index);
        orn_material g{};
        g.albedo[0] = 1.0f;  // Material.from default (gpu_structs.zig:45)
        g.albedo[1] = 0.0f;
        g.albedo[2] = 1.0f;
        g.albedo_texture_index = 0xffffffffu;
        if (m->texture) {
            textures.push_back(m->texture);
            m->texture->texture_id = static_cast<int64_t>(textures.size() - 1);
            g.albedo_texture_index = static_cast<uint32_t>(m->texture->texture_id);
        } else {
            std::memcpy(g.albedo, m->albedo.data(), 12);
        }
        g.fuzz = m->fuzz;
        g.ior = m->ior;
        g.type = m->type;
        materials.push_back(g);
        m->material_index = static_cast<int64_t>(materials.size() - 1);
        return static_cast<uint32_t>(m->material_index);
    }

    // bvh.zig:221-270
    orn_bvh_node build_blas(Triangle *t, size_t n)
    {
        if (n == 1) {
            orn_bvh_node node{};
            std::memcpy(node.lmin_or_v0, t[0].v0.data(), 12);
            std::memcpy(node.lmax_or_v1, t[0].v1.data(), 12);
            std::memcpy(node.rmin_or_v2, t[0].v2.data(), 12);
            node.left_or_custom_id = t[0].triangle_index;
            node.node_type = 3u;
            return node;
        }
        const size_t axis = rand.at_most(2);
        zig_heap_sort(t, n, [axis](const Triangle &a, const Triangle &b) { return a.aabb.min[axis] < b.aabb.min[axis]; });
        const size_t mid = n / 2;
        blas_nodes.push_back(build_blas(t, mid));
        const uint32_t left_id = static_cast<uint32_t>(blas_nodes.size() - 1);
        const Aabb left_box = bounding_box(t, t + mid, [](const Triangle &x) -> const Aabb & { return x.aabb; });
        blas_nodes.push_back(build_blas(t + mid, n - mid));
        const uint32_t right_id = static_cast<uint32_t>(blas_nodes.size() - 1);
        const Aabb right_box = bounding_box(t + mid, t + n, [](const Triangle &x) -> const Aabb & { return x.aabb; });
        return internal_node(left_box, left_id, right_box, right_id);
    }

    // bvh.zig:172-219
    void build_mesh(ornament_mesh *m)
    {
        const size_t ntri = m->vertex_indices.size() / 3;
        std::vector<Triangle> leafs(ntri);
        const uint32_t base = static_cast<uint32_t>(normal_indices.size() / 3);
        for (size_t i = 0; i < ntri; i++) {
            Triangle &t = leafs[i];
            t.v0 = m->vertices[m->vertex_indices[3 * i]];
            t.v1 = m->vertices[m->vertex_indices[3 * i + 1]];
            t.v2 = m->vertices[m->vertex_indices[3 * i + 2]];
            t.aabb.min = vmin(vmin(t.v0, t.v1), t.v2);
            t.aabb.max = vmax(vmax(t.v0, t.v1), t.v2);
            t.triangle_index = base + static_cast<uint32_t>(i);
        }
        const uint32_t nbase = static_cast<uint32_t>(normals.size());
        for (uint32_t idx : m->normal_indices) normal_indices.push_back(idx + nbase);
        for (const Vec &n : m->normals) normals.push_back({n[0], n[1], n[2], n[3]});
        const uint32_t ubase = static_cast<uint32_t>(uvs.size());
        for (uint32_t idx : m->uv_indices) uv_indices.push_back(idx + ubase);
        for (const auto &uv : m->uvs) uvs.push_back(uv);
        blas_nodes.push_back(build_blas(leafs.data(), ntri));
        m->bvh_id = static_cast<int64_t>(blas_nodes.size() - 1);
    }

    // bvh.zig:272-362
    orn_bvh_node build_tlas(Leaf *l, size_t n)
    {
        if (n == 1) {
            orn_bvh_node node{};
            const Aabb &b = l[0].aabb();
            std::memcpy(node.lmin_or_v0, b.min.data(), 12);
            std::memcpy(node.lmax_or_v1, b.max.data(), 12);
            Mat inv;
            switch (l[0].kind) {
            case Leaf::Sphere: {
                auto *s = static_cast<ornament_sphere *>(l[0].ptr);
                inverse(s->transform, inv);
                append_transform(inv);
                append_transform(s->transform);
                node.right_or_material_index = material_index(s->material);
                node.node_type = 1u;
                break;
            }
            case Leaf::Mesh: {
                auto *m = static_cast<ornament_mesh *>(l[0].ptr);
                inverse(m->transform, inv);
                append_transform(inv);
                append_transform(m->transform);
                node.left_or_custom_id = static_cast<uint32_t>(m->bvh_id);
                node.right_or_material_index = material_index(m->material);
                node.node_type = 2u;
                break;
            }
            default: {
                auto *mi = static_cast<ornament_mesh_instance *>(l[0].ptr);
                inverse(mi->transform, inv);
                append_transform(inv);
                append_transform(mi->transform);
                node.right_or_material_index = material_index(mi->material);
                node.node_type = 2u;
                node.left_or_custom_id = static_cast<uint32_t>(mi->mesh->bvh_id);
                break;
            }
            }
            node.transform_id = static_cast<uint32_t>(transforms.size() / 2 - 1);
            return node;
        }
        const size_t axis = rand.at_most(2);
        zig_heap_sort(l, n, [axis](const Leaf &a, const Leaf &b) { return a.aabb().min[axis] < b.aabb().min[axis]; });
        const size_t mid = n / 2;
        tlas_nodes.push_back(build_tlas(l, mid));
        const uint32_t left_id = static_cast<uint32_t>(tlas_nodes.size() - 1);
        const Aabb left_box = bounding_box(l, l + mid, [](const Leaf &x) -> const Aabb & { return x.aabb(); });
        tlas_nodes.push_back(build_tlas(l + mid, n - mid));
        const uint32_t right_id = static_cast<uint32_t>(tlas_nodes.size() - 1);
        const Aabb right_box = bounding_box(l + mid, l + n, [](const Leaf &x) -> const Aabb & { return x.aabb(); });
        return internal_node(left_box, left_id, right_box, right_id);
    }

    // bvh.zig:33-108
    orn_status build(ornament_scene *scene)
    {
        const size_t shapes = scene->spheres.size() + scene->meshes.size() + scene->mesh_instances.size();
        if (shapes == 0) return ORN_ERROR_EMPTY_SCENE;
        for (const auto &m : scene->meshes)
            if (m->vertex_indices.size() < 3) return ORN_ERROR_EMPTY_BVH;
        std::unordered_set<const ornament_mesh *> owned;
        for (const auto &m : scene->meshes) owned.insert(m.get());
        for (const auto &mi : scene->mesh_instances)
            if (!owned.count(mi->mesh)) return ORN_ERROR_INVALID_SCENE;  // `mesh.bvh_id orelse unreachable`
        std::vector<Leaf> leafs;
        leafs.reserve(shapes);
        for (const auto &s : scene->spheres) leafs.push_back({Leaf::Sphere, s.get()});
        for (const auto &mi : scene->mesh_instances) leafs.push_back({Leaf::MeshInstance, mi.get()});
        for (const auto &m : scene->meshes) {
            leafs.push_back({Leaf::Mesh, m.get()});
            build_mesh(m.get());
        }
        tlas_nodes.push_back(build_tlas(leafs.data(), leafs.size()));
        texture_descs.clear();
        for (ornament_texture *t : textures) {
            orn_texture_desc d{};
            d.width = t->width;
            d.height = t->height;
            d.num_components = t->num_components;
            d.bytes_per_component = t->bytes_per_component;
            d.is_hdr = t->is_hdr ? 1u : 0u;
            d.data = t->data.data();
            texture_descs.push_back(d);
            uint32_t src = 0;
            for (size_t i = 0; i < scene->textures.size(); i++)
                if (scene->textures[i].get() == t) src = static_cast<uint32_t>(i);
            texture_sources.push_back(src);
        }
        return ORN_OK;
    }

    void arrays(orn_scene_desc *d) const
    {
        d->tlas_nodes = tlas_nodes.data();
        d->tlas_nodes_len = tlas_nodes.size();
        d->blas_nodes = blas_nodes.data();
        d->blas_nodes_len = blas_nodes.size();
        d->normals = normals.empty() ? nullptr : normals[0].data();
        d->normals_len = normals.size();
        d->normal_indices = normal_indices.data();
        d->normal_indices_len = normal_indices.size();
        d->uvs = uvs.empty() ? nullptr : uvs[0].data();
        d->uvs_len = uvs.size();
        d->uv_indices = uv_indices.data();
        d->uv_indices_len = uv_indices.size();
        d->transforms = transforms.empty() ? nullptr : transforms[0].data();
        d->transforms_len = transforms.size();
        d->materials = materials.data();
        d->materials_len = materials.size();
        d->textures = texture_descs.data();
        d->textures_len = texture_descs.size();
    }
};

// state.zig:3-80
struct State {
    uint32_t width = 500, height = 500;
    uint32_t depth = 10;
    bool flip_y = false;
    float inverted_gamma = 1.0f;
    uint32_t iterations = 1;
    float ray_cast_epsilon = 0.001f;
    float current_iteration = 0.0f;
};

struct ornament_path_tracer {
    ornament_scene *scene = nullptr;
    State state;
    orn_ctx *ctx = nullptr;
    orn_device_scene *dscene = nullptr;
    orn_target *target = nullptr;
};

namespace {

orn_status get_or_create_target(ornament_path_tracer *pt, orn_target **out)
{
    if (!pt->target) {
        orn_status s = orn_target_create(pt->dscene, pt->state.width, pt->state.height, 0u, &pt->target);
        if (s != ORN_OK) return s;
    }
    *out = pt->target;
    return ORN_OK;
}

void camera_to_gpu(const Camera &c, orn_camera *g)
{
    std::memset(g, 0, sizeof(*g));
    std::memcpy(g->origin, c.origin.data(), 12);
    g->lens_radius = c.lens_radius;
    std::memcpy(g->lower_left_corner, c.lower_left_corner.data(), 12);
    std::memcpy(g->horizontal, c.horizontal.data(), 12);
    std::memcpy(g->vertical, c.vertical.data(), 12);
    std::memcpy(g->u, c.u.data(), 12);
    std::memcpy(g->v, c.v.data(), 12);
    std::memcpy(g->w, c.w.data(), 12);
}

void constant_params(const ornament_path_tracer *pt, orn_constant_params *p)
{
    std::memset(p, 0, sizeof(*p));
    camera_to_gpu(pt->scene->camera, &p->camera);
    p->depth = pt->state.depth;
    p->width = pt->state.width;
    p->height = pt->state.height;
    p->flip_y = pt->state.flip_y ? 1u : 0u;
    p->inverted_gamma = pt->state.inverted_gamma;
    p->ray_cast_epsilon = pt->state.ray_cast_epsilon;
    p->textures_count = static_cast<uint32_t>(pt->scene->textures.size());
    p->current_iteration = pt->state.current_iteration;
}

template <typename T>
orn_status add(std::vector<std::unique_ptr<T>> &list, std::unique_ptr<T> el, T **out)
{
    if (!out) return ORN_ERROR_INVALID_VALUE;
    *out = el.get();
    list.push_back(std::move(el));
    return ORN_OK;
}

orn_status new_material(ornament_scene *s, uint32_t type, const float *albedo, ornament_texture *tex,
                        float fuzz, float ior, ornament_material **out)
{
    if (!s) return ORN_ERROR_INVALID_HANDLE;
    if (!out || (!albedo && !tex && type != 2u)) return ORN_ERROR_INVALID_VALUE;
    auto m = std::make_unique<ornament_material>();
    if (albedo) m->albedo = load_vec(albedo);
    m->texture = tex;
    m->fuzz = fuzz;
    m->ior = ior;
    m->type = type;
    return add(s->materials, std::move(m), out);
}

orn_status new_mesh(ornament_scene *s, const std::vector<Vec> &vertices, const uint32_t *vi, size_t nvi,
                    const std::vector<Vec> &normals, const uint32_t *ni, size_t nni,
                    const float *uvs, size_t nuv, const uint32_t *ui, size_t nui, const Mat &transform,
                    ornament_material *material, ornament_mesh **out)
{
    if (!s) return ORN_ERROR_INVALID_HANDLE;
    if (!out || !material) return ORN_ERROR_INVALID_VALUE;
    for (size_t i = 0; i < nvi; i++)
        if (vi[i] >= vertices.size()) return ORN_ERROR_INVALID_VALUE;
    // scene.zig:274-287: bounds of the indexed triangles, then transformed
    const float inf = INFINITY;
    Aabb nt{vec(inf, inf, inf, 1.0f), vec(-inf, -inf, -inf, 1.0f)};
    for (size_t t = 0; t < nvi / 3; t++) {
        const Vec &a = vertices[vi[3 * t]], &b = vertices[vi[3 * t + 1]], &c = vertices[vi[3 * t + 2]];
        nt.min = vmin(nt.min, vmin(vmin(a, b), c));
        nt.max = vmax(nt.max, vmax(vmax(a, b), c));
    }
    auto m = std::make_unique<ornament_mesh>();
    m->transform = transform;
    m->material = material;
    m->not_transformed_aabb = nt;
    m->aabb = Aabb::transform(transform, nt);
    if (nuv == 0) {  // scene.zig:303-305
        m->uvs.assign(vertices.size(), {0.5f, 0.5f});
        m->uv_indices.assign(vi, vi + nvi);
    } else {
        for (size_t i = 0; i < nuv; i++) m->uvs.push_back({uvs[2 * i], uvs[2 * i + 1]});
        m->uv_indices.assign(ui, ui + nui);
    }
    m->vertices = vertices;
    m->vertex_indices.assign(vi, vi + nvi);
    m->normals = normals;
    m->normal_indices.assign(ni, ni + nni);
    return add(s->meshes, std::move(m), out);
}

}  // namespace

extern "C" {

orn_status ornament_scene_init(ornament_scene **out)
{
    if (!out) return ORN_ERROR_INVALID_VALUE;
    ornament_scene *s = new (std::nothrow) ornament_scene();
    if (!s) return ORN_ERROR_OUT_OF_MEMORY;
    // scene.zig:30-38
    s->camera = Camera::init(vec(1.0f, 1.0f, 1.0f, 1.0f), vec(0.0f, 0.0f, 0.0f, 1.0f),
                             vec(0.0f, 1.0f, 0.0f, 0.0f), 1.0f, 40.0f, 0.0f, 10.0f);
    *out = s;
    return ORN_OK;
}

void ornament_scene_deinit(ornament_scene *s) { delete s; }

orn_status ornament_scene_lambertian(ornament_scene *s, const float albedo[4], ornament_material **out)
{
    return new_material(s, 0u, albedo, nullptr, 0.0f, 0.0f, out);
}
orn_status ornament_scene_lambertian_texture(ornament_scene *s, ornament_texture *t, ornament_material **out)
{
    if (!t) return ORN_ERROR_INVALID_VALUE;
    return new_material(s, 0u, nullptr, t, 0.0f, 0.0f, out);
}
orn_status ornament_scene_metal(ornament_scene *s, const float albedo[4], float fuzz, ornament_material **out)
{
    return new_material(s, 1u, albedo, nullptr, fuzz, 0.0f, out);
}
orn_status ornament_scene_metal_texture(ornament_scene *s, ornament_texture *t, float fuzz,
                                        ornament_material **out)
{
    if (!t) return ORN_ERROR_INVALID_VALUE;
    return new_material(s, 1u, nullptr, t, fuzz, 0.0f, out);
}
orn_status ornament_scene_dielectric(ornament_scene *s, float ior, ornament_material **out)
{
    static const float zero[4] = {0.0f, 0.0f, 0.0f, 0.0f};  // `undefined` albedo in scene.zig:98
    return new_material(s, 2u, zero, nullptr, 0.0f, ior, out);
}
orn_status ornament_scene_diffuse_light(ornament_scene *s, const float albedo[4], ornament_material **out)
{
    return new_material(s, 3u, albedo, nullptr, 0.0f, 0.0f, out);
}
orn_status ornament_scene_diffuse_light_texture(ornament_scene *s, ornament_texture *t, ornament_material **out)
{
    if (!t) return ORN_ERROR_INVALID_VALUE;
    return new_material(s, 3u, nullptr, t, 0.0f, 0.0f, out);
}

orn_status ornament_scene_create_sphere(ornament_scene *s, const float center[4], float radius,
                                        ornament_material *m, ornament_sphere **out)
{
    if (!s) return ORN_ERROR_INVALID_HANDLE;
    if (!center || !m) return ORN_ERROR_INVALID_VALUE;
    auto sp = std::make_unique<ornament_sphere>();
    const Vec c = load_vec(center);
    const Vec r = vec(radius, radius, radius, 0.0f);
    sp->transform = mul(scaling(radius, radius, radius), translation(c[0], c[1], c[2]));
    sp->material = m;
    sp->aabb = Aabb{sub(c, r), Vec{c[0] + r[0], c[1] + r[1], c[2] + r[2], c[3] + r[3]}};
    return add(s->spheres, std::move(sp), out);
}

// scene.zig:140-233: a 60 x 30 UV sphere as a mesh
orn_status ornament_scene_create_sphere_mesh(ornament_scene *s, const float center[4], float radius,
                                             ornament_material *m, ornament_mesh **out)
{
    if (!s) return ORN_ERROR_INVALID_HANDLE;
    if (!center || !m || !out) return ORN_ERROR_INVALID_VALUE;
    std::vector<Vec> vertices, normals;
    std::vector<float> uvs;
    std::vector<uint32_t> indices;
    const float facing = 1.0f;
    const uint32_t hs = 60, vs = 30;
    auto push_uv = [&](const Vec &n) {
        const auto uv = sphere_tex_coord(n);
        uvs.push_back(uv[0]);
        uvs.push_back(uv[1]);
    };
    vertices.push_back(vec(0.0f, 1.0f, 0.0f, 1.0f));
    normals.push_back(vec(0.0f, facing, 0.0f, 0.0f));
    push_uv(normals.back());
    for (uint32_t v = 1; v < vs; v++) {
        const float theta = static_cast<float>(v) / static_cast<float>(vs) * kPi;
        const float sin_theta = static_cast<float>(std::sin(static_cast<double>(theta)));
        for (uint32_t h = 0; h < hs; h++) {
            const float phi = static_cast<float>(h) / static_cast<float>(hs) * kPi * 2.0f;
            const float x = sin_theta * static_cast<float>(std::sin(static_cast<double>(phi))) * 1.0f;
            const float z = sin_theta * static_cast<float>(std::cos(static_cast<double>(phi))) * 1.0f;
            const float y = static_cast<float>(std::cos(static_cast<double>(theta))) * 1.0f;
            vertices.push_back(vec(x, y, z, 1.0f));
            normals.push_back(normalize3(scale(facing, vec(x, y, z, 0.0f))));
            push_uv(normals.back());
            if (v == 1) {
                indices.push_back(0);
                indices.push_back(h + 1);
                indices.push_back(h < hs - 1 ? h + 2 : 1);
            } else {
                const uint32_t i = h + ((v - 1) * hs) + 1, j = i - hs;
                const uint32_t k = h < hs - 1 ? j + 1 : j - (hs - 1);
                const uint32_t l = h < hs - 1 ? i + 1 : i - (hs - 1);
                for (uint32_t q : {j, i, k, k, i, l}) indices.push_back(q);
            }
        }
    }
    vertices.push_back(vec(0.0f, -1.0f, 0.0f, 1.0f));
    normals.push_back(vec(0.0f, -facing, 0.0f, 0.0f));
    push_uv(normals.back());
    const uint32_t end = static_cast<uint32_t>(vertices.size()) - 1;
    for (uint32_t h = 0; h < hs; h++) {
        const uint32_t i = end - hs + h;
        indices.push_back(i);
        indices.push_back(end);
        indices.push_back(h < hs - 1 ? i + 1 : end - hs);
    }
    const Vec c = load_vec(center);
    const Mat t = mul(scaling(radius, radius, radius), translation(c[0], c[1], c[2]));
    return new_mesh(s, vertices, indices.data(), indices.size(), normals, indices.data(), indices.size(),
                    uvs.data(), uvs.size() / 2, indices.data(), indices.size(), t, m, out);
}

orn_status ornament_scene_create_plane_mesh(ornament_scene *s, const float center[4], float side1,
                                            float side2, const float normal[4], ornament_material *m,
                                            ornament_mesh **out)
{
    if (!s) return ORN_ERROR_INVALID_HANDLE;
    if (!center || !normal) return ORN_ERROR_INVALID_VALUE;
    const std::vector<Vec> vertices = {vec(-0.5f, 0.0f, -0.5f, 1.0f), vec(-0.5f, 0.0f, 0.5f, 1.0f),
                                       vec(0.5f, 0.0f, 0.5f, 1.0f), vec(0.5f, 0.0f, -0.5f, 1.0f)};
    static const uint32_t indices[6] = {3, 1, 0, 2, 1, 3};
    const Vec unit_y = vec(0.0f, 1.0f, 0.0f, 0.0f);
    const std::vector<Vec> normals = {unit_y, unit_y, unit_y, unit_y};
    const Mat rotation = rotation_between_vectors(load_vec(normal), unit_y);
    const Vec c = load_vec(center);
    const Mat t = mul(mul(scaling(side1, 1.0f, side2), rotation), translation(c[0], c[1], c[2]));
    return new_mesh(s, vertices, indices, 6, normals, indices, 6, nullptr, 0, nullptr, 0, t, m, out);
}

orn_status ornament_scene_create_mesh(ornament_scene *s, const float *vertices, size_t nv,
                                      const uint32_t *vi, size_t nvi, const float *normals, size_t nn,
                                      const uint32_t *ni, size_t nni, const float *uvs, size_t nuv,
                                      const uint32_t *ui, size_t nui, const float transform[16],
                                      ornament_material *m, ornament_mesh **out)
{
    if (!s) return ORN_ERROR_INVALID_HANDLE;
    if ((!vertices && nv) || (!vi && nvi) || (!normals && nn) || (!ni && nni) || (!uvs && nuv) ||
        (!ui && nui) || !transform)
        return ORN_ERROR_INVALID_VALUE;
    std::vector<Vec> v(nv), n(nn);
    for (size_t i = 0; i < nv; i++) v[i] = load_vec(vertices + 4 * i);
    for (size_t i = 0; i < nn; i++) n[i] = load_vec(normals + 4 * i);
    Mat t;
    std::memcpy(t.data(), transform, 64);
    return new_mesh(s, v, vi, nvi, n, ni, nni, uvs, nuv, ui, nui, t, m, out);
}

orn_status ornament_scene_create_mesh_instance(ornament_scene *s, ornament_mesh *mesh, const float transform[16],
                                               ornament_material *m, ornament_mesh_instance **out)
{
    if (!s) return ORN_ERROR_INVALID_HANDLE;
    if (!mesh || !transform || !m) return ORN_ERROR_INVALID_VALUE;
    auto mi = std::make_unique<ornament_mesh_instance>();
    mi->mesh = mesh;
    mi->material = m;
    std::memcpy(mi->transform.data(), transform, 64);
    mi->aabb = Aabb::transform(mi->transform, mesh->not_transformed_aabb);
    return add(s->mesh_instances, std::move(mi), out);
}

orn_status ornament_scene_create_texture(ornament_scene *s, const uint8_t *data, size_t data_len, uint32_t width,
                                         uint32_t height, uint32_t num_components, uint32_t bytes_per_component,
                                         int is_hdr, float gamma, ornament_texture **out)
{
    if (!s) return ORN_ERROR_INVALID_HANDLE;
    // bounded sizes keep every product below 2^32 (8192 * 8192 * 4 * 8)
    if (width > ORN_MAX_TEXTURE_DIMENSION || height > ORN_MAX_TEXTURE_DIMENSION || num_components > 4u ||
        bytes_per_component > 8u)
        return ORN_ERROR_UNSUPPORTED_TEXTURE;
    const size_t bytes_per_row = size_t(width) * num_components * bytes_per_component;
    if (!data || !out || data_len < bytes_per_row * height) return ORN_ERROR_INVALID_VALUE;
    auto t = std::make_unique<ornament_texture>();
    t->data.assign(data, data + bytes_per_row * height);
    t->width = width;
    t->height = height;
    t->num_components = num_components;
    t->bytes_per_component = bytes_per_component;
    t->bytes_per_row = static_cast<uint32_t>(bytes_per_row);
    t->is_hdr = is_hdr != 0;
    t->gamma = gamma;
    return add(s->textures, std::move(t), out);
}

orn_status ornament_scene_attach_sphere(ornament_scene *s, ornament_sphere *x)
{
    if (!s || !x) return ORN_ERROR_INVALID_HANDLE;
    s->attached_spheres.push_back(x);
    return ORN_OK;
}
orn_status ornament_scene_attach_mesh(ornament_scene *s, ornament_mesh *x)
{
    if (!s || !x) return ORN_ERROR_INVALID_HANDLE;
    s->attached_meshes.push_back(x);
    return ORN_OK;
}
orn_status ornament_scene_attach_mesh_instance(ornament_scene *s, ornament_mesh_instance *x)
{
    if (!s || !x) return ORN_ERROR_INVALID_HANDLE;
    s->attached_mesh_instances.push_back(x);
    return ORN_OK;
}

orn_status ornament_scene_counts(const ornament_scene *s, uint32_t counts[5])
{
    if (!s) return ORN_ERROR_INVALID_HANDLE;
    if (!counts) return ORN_ERROR_INVALID_VALUE;
    counts[0] = static_cast<uint32_t>(s->spheres.size());
    counts[1] = static_cast<uint32_t>(s->meshes.size());
    counts[2] = static_cast<uint32_t>(s->mesh_instances.size());
    counts[3] = static_cast<uint32_t>(s->materials.size());
    counts[4] = static_cast<uint32_t>(s->textures.size());
    return ORN_OK;
}

orn_status ornament_scene_set_camera(ornament_scene *s, const float lookfrom[4], const float lookat[4],
                                     const float vup[4], float aspect_ratio, float vfov, float aperture,
                                     float focus_dist)
{
    if (!s) return ORN_ERROR_INVALID_HANDLE;
    if (!lookfrom || !lookat || !vup) return ORN_ERROR_INVALID_VALUE;
    s->camera = Camera::init(load_vec(lookfrom), load_vec(lookat), load_vec(vup), aspect_ratio, vfov, aperture,
                             focus_dist);
    return ORN_OK;
}

orn_status ornament_camera_set_look_at(ornament_scene *s, const float lookfrom[4], const float lookat[4],
                                       const float vup[4])
{
    if (!s) return ORN_ERROR_INVALID_HANDLE;
    if (!lookfrom || !lookat || !vup) return ORN_ERROR_INVALID_VALUE;
    const Camera &c = s->camera;
    s->camera = Camera::init(load_vec(lookfrom), load_vec(lookat), load_vec(vup), c.aspect_ratio, c.vfov,
                             2.0f * c.lens_radius, c.focus_dist);
    return ORN_OK;
}

orn_status ornament_camera_set_aspect_ratio(ornament_scene *s, float aspect_ratio)
{
    if (!s) return ORN_ERROR_INVALID_HANDLE;
    const Camera &c = s->camera;
    s->camera = Camera::init(c.lookfrom, c.lookat, c.vup, aspect_ratio, c.vfov, 2.0f * c.lens_radius, c.focus_dist);
    return ORN_OK;
}

orn_status ornament_camera_get(const ornament_scene *s, orn_camera *out)
{
    if (!s) return ORN_ERROR_INVALID_HANDLE;
    if (!out) return ORN_ERROR_INVALID_VALUE;
    camera_to_gpu(s->camera, out);
    return ORN_OK;
}

orn_status ornament_camera_get_dirty(const ornament_scene *s, int *dirty)
{
    if (!s) return ORN_ERROR_INVALID_HANDLE;
    if (!dirty) return ORN_ERROR_INVALID_VALUE;
    *dirty = s->camera.dirty ? 1 : 0;
    return ORN_OK;
}

orn_status ornament_bvh_init(ornament_scene *s, int row_major, ornament_bvh **out)
{
    if (!s) return ORN_ERROR_INVALID_HANDLE;
    if (!out) return ORN_ERROR_INVALID_VALUE;
    *out = nullptr;
    auto b = std::make_unique<ornament_bvh>();
    b->row_major_transforms = row_major != 0;
    // with the shared stream, builds run one at a time and each takes the
    // generator where the previous one left it
    std::unique_lock<std::mutex> lock(g_bvh_prng.mutex);
    if (g_bvh_prng.shared)
        std::memcpy(b->rand.s, g_bvh_prng.state, sizeof(b->rand.s));
    else
        lock.unlock();
    orn_status st = b->build(s);
    if (lock.owns_lock()) std::memcpy(g_bvh_prng.state, b->rand.s, sizeof(b->rand.s));
    if (st != ORN_OK) return st;
    *out = b.release();
    return ORN_OK;
}

orn_status ornament_bvh_shared_generator(int mode)
{
    if (mode != 0 && mode != 1) return ORN_ERROR_INVALID_VALUE;
    std::lock_guard<std::mutex> lock(g_bvh_prng.mutex);
    g_bvh_prng.shared = mode == 1;
    const Xoshiro256 fresh{1244u};  // bvh.zig:12: DefaultPrng.init(1244)
    std::memcpy(g_bvh_prng.state, fresh.s, sizeof(g_bvh_prng.state));
    return ORN_OK;
}

void ornament_bvh_deinit(ornament_bvh *b) { delete b; }

orn_status ornament_bvh_arrays(const ornament_bvh *b, orn_scene_desc *out)
{
    if (!b) return ORN_ERROR_INVALID_HANDLE;
    if (!out) return ORN_ERROR_INVALID_VALUE;
    b->arrays(out);
    return ORN_OK;
}

orn_status ornament_bvh_texture_sources(const ornament_bvh *b, uint32_t *dst, size_t len)
{
    if (!b) return ORN_ERROR_INVALID_HANDLE;
    if (len < b->texture_sources.size() || (!dst && len)) return ORN_ERROR_SIZE_MISMATCH;
    for (size_t i = 0; i < b->texture_sources.size(); i++) dst[i] = b->texture_sources[i];
    return ORN_OK;
}

orn_status ornament_path_tracer_init(int device, ornament_scene *scene, ornament_path_tracer **out)
{
    if (!scene) return ORN_ERROR_INVALID_HANDLE;
    if (!out) return ORN_ERROR_INVALID_VALUE;
    *out = nullptr;
    auto pt = std::make_unique<ornament_path_tracer>();
    orn_status st = orn_device_open(device, &pt->ctx);
    if (st != ORN_OK) return st;
    ornament_bvh *bvh = nullptr;
    st = ornament_bvh_init(scene, 0, &bvh);
    if (st == ORN_OK) {
        orn_scene_desc d;
        bvh->arrays(&d);
        st = orn_scene_upload(pt->ctx, &d, &pt->dscene);
        ornament_bvh_deinit(bvh);  // `defer bvh.deinit()` (path_tracer.zig:56)
    }
    if (st != ORN_OK) {
        orn_device_close(pt->ctx);
        return st;
    }
    pt->scene = scene;  // ownership moves to the path tracer
    *out = pt.release();
    return ORN_OK;
}

orn_status ornament_path_tracer_deinit(ornament_path_tracer *pt)
{
    if (!pt) return ORN_ERROR_INVALID_HANDLE;
    if (pt->target) orn_target_destroy(pt->target);
    if (pt->dscene) orn_scene_destroy(pt->dscene);
    if (pt->ctx) orn_device_close(pt->ctx);
    delete pt->scene;
    delete pt;
    return ORN_OK;
}

ornament_scene *ornament_path_tracer_scene(ornament_path_tracer *pt) { return pt ? pt->scene : nullptr; }

// path_tracer.zig:198-210: a new size drops the target; state is kept
orn_status ornament_path_tracer_set_resolution(ornament_path_tracer *pt, uint32_t width, uint32_t height)
{
    if (!pt) return ORN_ERROR_INVALID_HANDLE;
    if (width == 0 || height == 0) return ORN_ERROR_INVALID_VALUE;
    if (pt->state.width != width || pt->state.height != height) {
        pt->state.width = width;
        pt->state.height = height;
        if (pt->target) {
            orn_target_destroy(pt->target);
            pt->target = nullptr;
        }
    }
    return ORN_OK;
}

// path_tracer.zig:212-265: update() per iteration, then the dispatches. The
// camera cannot change inside one render() call, so the dirty check of the
// first update() is the only one that can fire; all iterations go to the GPU
// in one launch and current_iteration advances as the per-iteration updates
// would advance it.
orn_status ornament_path_tracer_render(ornament_path_tracer *pt)
{
    if (!pt) return ORN_ERROR_INVALID_HANDLE;
    orn_target *t = nullptr;
    orn_status st = get_or_create_target(pt, &t);
    if (st != ORN_OK) return st;
    if (pt->scene->camera.dirty) {
        pt->scene->camera.dirty = false;
        pt->state.current_iteration = 0.0f;  // State.reset
    }
    orn_constant_params p;
    constant_params(pt, &p);
    st = orn_render(t, &p, pt->state.iterations);
    if (st != ORN_OK) return st;
    const uint32_t n = pt->state.iterations > 1 ? pt->state.iterations : 1;
    for (uint32_t i = 0; i < n; i++) pt->state.current_iteration += 1.0f;
    return ORN_OK;
}

orn_status ornament_path_tracer_get_framebuffer(ornament_path_tracer *pt, float *dst, size_t pixels)
{
    if (!pt) return ORN_ERROR_INVALID_HANDLE;
    orn_target *t = nullptr;
    orn_status st = get_or_create_target(pt, &t);
    if (st != ORN_OK) return st;
    return orn_get_framebuffer(t, dst, pixels);
}

orn_status ornament_path_tracer_get_accumulation(ornament_path_tracer *pt, float *dst, size_t pixels)
{
    if (!pt) return ORN_ERROR_INVALID_HANDLE;
    orn_target *t = nullptr;
    orn_status st = get_or_create_target(pt, &t);
    if (st != ORN_OK) return st;
    return orn_get_accumulation(t, dst, pixels);
}

orn_status ornament_path_tracer_last_render_ms(ornament_path_tracer *pt, float *ms)
{
    if (!pt) return ORN_ERROR_INVALID_HANDLE;
    if (!pt->target) {
        if (ms) *ms = 0.0f;
        return ms ? ORN_OK : ORN_ERROR_INVALID_VALUE;
    }
    return orn_target_last_render_ms(pt->target, ms);
}

#define ORN_STATE_CHECK(pt, out)             \
    if (!(pt)) return ORN_ERROR_INVALID_HANDLE; \
    if (!(out)) return ORN_ERROR_INVALID_VALUE

orn_status ornament_state_set_depth(ornament_path_tracer *pt, uint32_t v)
{
    if (!pt) return ORN_ERROR_INVALID_HANDLE;
    pt->state.depth = v;
    return ORN_OK;
}
orn_status ornament_state_get_depth(ornament_path_tracer *pt, uint32_t *v)
{
    ORN_STATE_CHECK(pt, v);
    *v = pt->state.depth;
    return ORN_OK;
}
orn_status ornament_state_set_iterations(ornament_path_tracer *pt, uint32_t v)
{
    if (!pt) return ORN_ERROR_INVALID_HANDLE;
    pt->state.iterations = v;
    return ORN_OK;
}
orn_status ornament_state_get_iterations(ornament_path_tracer *pt, uint32_t *v)
{
    ORN_STATE_CHECK(pt, v);
    *v = pt->state.iterations;
    return ORN_OK;
}
orn_status ornament_state_set_gamma(ornament_path_tracer *pt, float gamma)
{
    if (!pt) return ORN_ERROR_INVALID_HANDLE;
    pt->state.inverted_gamma = 1.0f / gamma;  // state.zig:35-37
    return ORN_OK;
}
orn_status ornament_state_get_gamma(ornament_path_tracer *pt, float *v)
{
    ORN_STATE_CHECK(pt, v);
    *v = 1.0f / pt->state.inverted_gamma;
    return ORN_OK;
}
orn_status ornament_state_set_flip_y(ornament_path_tracer *pt, int v)
{
    if (!pt) return ORN_ERROR_INVALID_HANDLE;
    pt->state.flip_y = v != 0;
    return ORN_OK;
}
orn_status ornament_state_get_flip_y(ornament_path_tracer *pt, int *v)
{
    ORN_STATE_CHECK(pt, v);
    *v = pt->state.flip_y ? 1 : 0;
    return ORN_OK;
}
orn_status ornament_state_set_ray_cast_epsilon(ornament_path_tracer *pt, float v)
{
    if (!pt) return ORN_ERROR_INVALID_HANDLE;
    pt->state.ray_cast_epsilon = v;
    return ORN_OK;
}
orn_status ornament_state_get_ray_cast_epsilon(ornament_path_tracer *pt, float *v)
{
    ORN_STATE_CHECK(pt, v);
    *v = pt->state.ray_cast_epsilon;
    return ORN_OK;
}
orn_status ornament_state_get_resolution(ornament_path_tracer *pt, uint32_t *w, uint32_t *h)
{
    ORN_STATE_CHECK(pt, w);
    if (!h) return ORN_ERROR_INVALID_VALUE;
    *w = pt->state.width;
    *h = pt->state.height;
    return ORN_OK;
}
orn_status ornament_state_get_current_iteration(ornament_path_tracer *pt, float *v)
{
    ORN_STATE_CHECK(pt, v);
    *v = pt->state.current_iteration;
    return ORN_OK;
}
orn_status ornament_path_tracer_constant_params(ornament_path_tracer *pt, orn_constant_params *out)
{
    ORN_STATE_CHECK(pt, out);
    constant_params(pt, out);
    return ORN_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Scene files and images
// ---------------------------------------------------------------------------
namespace {

constexpr char kSceneMagic[8] = {'O', 'R', 'N', 'S', 'C', 'E', 'N', 'E'};
constexpr uint32_t kSceneVersion = 1;
constexpr uint64_t kMaxArray = 1ull << 32;  // elements of one array in a scene file

// Little-endian writer/reader over a byte vector (x86-64 and the GPU host
// are little-endian, so values are copied as they are).
struct Writer {
    std::vector<uint8_t> b;
    template <typename T> void put(const T &v)
    {
        const uint8_t *p = reinterpret_cast<const uint8_t *>(&v);
        b.insert(b.end(), p, p + sizeof(T));
    }
    void put_bytes(const void *p, size_t n)
    {
        const uint8_t *q = static_cast<const uint8_t *>(p);
        b.insert(b.end(), q, q + n);
    }
};

struct Reader {
    const std::vector<uint8_t> &b;
    size_t at = 0;
    bool ok = true;
    template <typename T> T get()
    {
        T v{};
        if (!ok || b.size() - at < sizeof(T)) {
            ok = false;
            return v;
        }
        std::memcpy(&v, b.data() + at, sizeof(T));
        at += sizeof(T);
        return v;
    }
    // n elements of T; fails on sizes the file cannot hold
    template <typename T> bool get_array(uint64_t n, std::vector<T> &out)
    {
        if (!ok || n > kMaxArray || n > (b.size() - at) / sizeof(T)) return ok = false;
        out.resize(n);
        if (n) std::memcpy(out.data(), b.data() + at, n * sizeof(T));
        at += n * sizeof(T);
        return true;
    }
};

bool read_file(const char *path, std::vector<uint8_t> &out)
{
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    uint8_t buf[65536];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) out.insert(out.end(), buf, buf + n);
    const bool err = std::ferror(f) != 0;
    std::fclose(f);
    return !err;
}

bool write_file(const char *path, const std::vector<uint8_t> &data)
{
    FILE *f = std::fopen(path, "wb");
    if (!f) return false;
    const bool ok = std::fwrite(data.data(), 1, data.size(), f) == data.size();
    return (std::fclose(f) == 0) && ok;
}

template <typename P> uint32_t index_of(const std::vector<std::unique_ptr<P>> &list, const P *p)
{
    for (size_t i = 0; i < list.size(); i++)
        if (list[i].get() == p) return static_cast<uint32_t>(i);
    return 0xffffffffu;
}

template <typename P> bool attached(const std::vector<P *> &list, const P *p)
{
    for (const P *q : list)
        if (q == p) return true;
    return false;
}

}  // namespace

extern "C" {

orn_status ornament_scene_save(const ornament_scene *s, const ornament_render_settings *settings, const char *path)
{
    if (!s) return ORN_ERROR_INVALID_HANDLE;
    if (!path) return ORN_ERROR_INVALID_VALUE;
    Writer w;
    w.put_bytes(kSceneMagic, 8);
    w.put(kSceneVersion);
    w.put(static_cast<uint32_t>(settings ? 1u : 0u));
    const Camera &c = s->camera;
    for (int i = 0; i < 3; i++) w.put(c.lookfrom[i]);
    for (int i = 0; i < 3; i++) w.put(c.lookat[i]);
    for (int i = 0; i < 3; i++) w.put(c.vup[i]);
    w.put(c.aspect_ratio);
    w.put(c.vfov);
    w.put(c.lens_radius * 2.0f);  // aperture (camera.zig:60)
    w.put(c.focus_dist);
    if (settings) w.put(*settings);
    w.put(static_cast<uint32_t>(s->textures.size()));
    w.put(static_cast<uint32_t>(s->materials.size()));
    w.put(static_cast<uint32_t>(s->spheres.size()));
    w.put(static_cast<uint32_t>(s->meshes.size()));
    w.put(static_cast<uint32_t>(s->mesh_instances.size()));
    for (const auto &t : s->textures) {
        w.put(t->width);
        w.put(t->height);
        w.put(t->num_components);
        w.put(t->bytes_per_component);
        w.put(static_cast<uint32_t>(t->is_hdr ? 1u : 0u));
        w.put(t->gamma);
        w.put(static_cast<uint64_t>(t->data.size()));
        w.put_bytes(t->data.data(), t->data.size());
        while (w.b.size() % 4) w.b.push_back(0);
    }
    for (const auto &m : s->materials) {
        w.put(m->type);
        w.put(m->texture ? index_of(s->textures, m->texture) : 0xffffffffu);
        for (int i = 0; i < 3; i++) w.put(m->albedo[i]);
        w.put(m->fuzz);
        w.put(m->ior);
    }
    for (const auto &sp : s->spheres) {
        // S(r) * T(c) in row-vector form: r on the diagonal, c in the last row
        w.put(sp->transform[12]);
        w.put(sp->transform[13]);
        w.put(sp->transform[14]);
        w.put(sp->transform[0]);
        w.put(index_of(s->materials, sp->material));
        w.put(static_cast<uint32_t>(attached(s->attached_spheres, sp.get())));
    }
    for (const auto &m : s->meshes) {
        w.put(static_cast<uint64_t>(m->vertices.size()));
        w.put(static_cast<uint64_t>(m->vertex_indices.size()));
        w.put(static_cast<uint64_t>(m->normals.size()));
        w.put(static_cast<uint64_t>(m->normal_indices.size()));
        w.put(static_cast<uint64_t>(m->uvs.size()));
        w.put(static_cast<uint64_t>(m->uv_indices.size()));
        for (const Vec &v : m->vertices) w.put_bytes(v.data(), 16);
        w.put_bytes(m->vertex_indices.data(), 4 * m->vertex_indices.size());
        for (const Vec &v : m->normals) w.put_bytes(v.data(), 16);
        w.put_bytes(m->normal_indices.data(), 4 * m->normal_indices.size());
        for (const auto &uv : m->uvs) w.put_bytes(uv.data(), 8);
        w.put_bytes(m->uv_indices.data(), 4 * m->uv_indices.size());
        w.put_bytes(m->transform.data(), 64);
        w.put(index_of(s->materials, m->material));
        w.put(static_cast<uint32_t>(attached(s->attached_meshes, m.get())));
    }
    for (const auto &mi : s->mesh_instances) {
        w.put(index_of(s->meshes, mi->mesh));
        w.put_bytes(mi->transform.data(), 64);
        w.put(index_of(s->materials, mi->material));
        w.put(static_cast<uint32_t>(attached(s->attached_mesh_instances, mi.get())));
    }
    return write_file(path, w.b) ? ORN_OK : ORN_ERROR_IO;
}

orn_status ornament_scene_load(const char *path, ornament_scene **out, ornament_render_settings *settings,
                               int *has_settings)
{
    if (!path || !out) return ORN_ERROR_INVALID_VALUE;
    *out = nullptr;
    std::vector<uint8_t> bytes;
    if (!read_file(path, bytes)) return ORN_ERROR_IO;
    Reader r{bytes};
    char magic[8];
    for (char &ch : magic) ch = static_cast<char>(r.get<uint8_t>());
    if (!r.ok || std::memcmp(magic, kSceneMagic, 8) != 0 || r.get<uint32_t>() != kSceneVersion)
        return ORN_ERROR_INVALID_FILE;
    const uint32_t flags = r.get<uint32_t>();
    float cam[13];
    for (float &v : cam) v = r.get<float>();
    ornament_render_settings rs{};
    if (flags & 1u) rs = r.get<ornament_render_settings>();
    uint32_t counts[5];
    for (uint32_t &n : counts) n = r.get<uint32_t>();
    if (!r.ok) return ORN_ERROR_INVALID_FILE;
    ornament_scene *s = nullptr;
    orn_status st = ornament_scene_init(&s);
    if (st != ORN_OK) return st;
    std::unique_ptr<ornament_scene> guard(s);
    auto bad = [] { return ORN_ERROR_INVALID_FILE; };
    std::vector<ornament_texture *> tex;
    for (uint32_t i = 0; i < counts[0]; i++) {
        const uint32_t wdt = r.get<uint32_t>(), hgt = r.get<uint32_t>(), comps = r.get<uint32_t>(),
                       bpc = r.get<uint32_t>(), hdr = r.get<uint32_t>();
        const float gamma = r.get<float>();
        const uint64_t n = r.get<uint64_t>();
        std::vector<uint8_t> data;
        if (!r.get_array(n, data)) return bad();
        while (r.at % 4 && r.ok) (void)r.get<uint8_t>();
        ornament_texture *t = nullptr;
        if (wdt > ORN_MAX_TEXTURE_DIMENSION || hgt > ORN_MAX_TEXTURE_DIMENSION || comps > 4u || bpc > 8u ||
            uint64_t(wdt) * hgt * comps * bpc != n ||
            ornament_scene_create_texture(s, data.data(), data.size(), wdt, hgt, comps, bpc, int(hdr), gamma, &t))
            return bad();
        tex.push_back(t);
    }
    std::vector<ornament_material *> mats;
    for (uint32_t i = 0; i < counts[1]; i++) {
        const uint32_t type = r.get<uint32_t>(), ti = r.get<uint32_t>();
        float albedo[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int k = 0; k < 3; k++) albedo[k] = r.get<float>();
        const float fuzz = r.get<float>(), ior = r.get<float>();
        if (!r.ok || type > 3u || (ti != 0xffffffffu && ti >= tex.size())) return bad();
        ornament_texture *t = ti == 0xffffffffu ? nullptr : tex[ti];
        ornament_material *m = nullptr;
        switch (type) {
        case 0: st = t ? ornament_scene_lambertian_texture(s, t, &m) : ornament_scene_lambertian(s, albedo, &m); break;
        case 1: st = t ? ornament_scene_metal_texture(s, t, fuzz, &m) : ornament_scene_metal(s, albedo, fuzz, &m); break;
        case 2: st = ornament_scene_dielectric(s, ior, &m); break;
        default: st = t ? ornament_scene_diffuse_light_texture(s, t, &m) : ornament_scene_diffuse_light(s, albedo, &m);
        }
        if (st != ORN_OK) return st;
        mats.push_back(m);
    }
    auto material = [&](uint32_t i) { return i < mats.size() ? mats[i] : nullptr; };
    for (uint32_t i = 0; i < counts[2]; i++) {
        float c[4] = {0.0f, 0.0f, 0.0f, 1.0f};
        for (int k = 0; k < 3; k++) c[k] = r.get<float>();
        const float radius = r.get<float>();
        ornament_material *m = material(r.get<uint32_t>());
        const uint32_t att = r.get<uint32_t>();
        ornament_sphere *sp = nullptr;
        if (!r.ok || !m || ornament_scene_create_sphere(s, c, radius, m, &sp)) return bad();
        if (att) (void)ornament_scene_attach_sphere(s, sp);
    }
    std::vector<ornament_mesh *> meshes;
    for (uint32_t i = 0; i < counts[3]; i++) {
        uint64_t n[6];
        for (uint64_t &v : n) v = r.get<uint64_t>();
        std::vector<float> v, nn, uv;
        std::vector<uint32_t> vi, ni, ui;
        if (!r.get_array(n[0] * 4, v) || !r.get_array(n[1], vi) || !r.get_array(n[2] * 4, nn) ||
            !r.get_array(n[3], ni) || !r.get_array(n[4] * 2, uv) || !r.get_array(n[5], ui))
            return bad();
        float t[16];
        for (float &x : t) x = r.get<float>();
        ornament_material *m = material(r.get<uint32_t>());
        const uint32_t att = r.get<uint32_t>();
        // indices must stay inside their arrays (the BVH build reads them)
        for (uint32_t x : ni)
            if (x >= n[2]) return bad();
        for (uint32_t x : ui)
            if (x >= n[4]) return bad();
        if (n[1] % 3 || n[3] < n[1] || n[5] < n[1]) return bad();
        ornament_mesh *mesh = nullptr;
        if (!r.ok || !m ||
            ornament_scene_create_mesh(s, v.data(), n[0], vi.data(), n[1], nn.data(), n[2], ni.data(), n[3],
                                       uv.data(), n[4], ui.data(), n[5], t, m, &mesh))
            return bad();
        if (att) (void)ornament_scene_attach_mesh(s, mesh);
        meshes.push_back(mesh);
    }
    for (uint32_t i = 0; i < counts[4]; i++) {
        const uint32_t mi = r.get<uint32_t>();
        float t[16];
        for (float &x : t) x = r.get<float>();
        ornament_material *m = material(r.get<uint32_t>());
        const uint32_t att = r.get<uint32_t>();
        ornament_mesh_instance *inst = nullptr;
        if (!r.ok || !m || mi >= meshes.size() ||
            ornament_scene_create_mesh_instance(s, meshes[mi], t, m, &inst))
            return bad();
        if (att) (void)ornament_scene_attach_mesh_instance(s, inst);
    }
    if (!r.ok || r.at != bytes.size()) return ORN_ERROR_INVALID_FILE;
    const float from[4] = {cam[0], cam[1], cam[2], 1.0f}, at[4] = {cam[3], cam[4], cam[5], 1.0f},
                up[4] = {cam[6], cam[7], cam[8], 0.0f};
    st = ornament_scene_set_camera(s, from, at, up, cam[9], cam[10], cam[11], cam[12]);
    if (st != ORN_OK) return st;
    if (settings) *settings = rs;
    if (has_settings) *has_settings = (flags & 1u) ? 1 : 0;
    *out = guard.release();
    return ORN_OK;
}

orn_status ornament_image_write_pfm(const char *path, const float *rgba, uint32_t width, uint32_t height)
{
    if (!path || !rgba || width == 0 || height == 0) return ORN_ERROR_INVALID_VALUE;
    Writer w;
    char head[64];
    const int n = std::snprintf(head, sizeof(head), "PF\n%u %u\n-1.0\n", width, height);
    w.put_bytes(head, static_cast<size_t>(n));
    // a negative scale means little-endian; rows run bottom to top, as row 0
    // of the framebuffer is the bottom of the picture
    const size_t px = size_t(width) * height;
    for (size_t i = 0; i < px; i++) w.put_bytes(rgba + 4 * i, 12);
    return write_file(path, w.b) ? ORN_OK : ORN_ERROR_IO;
}

orn_status ornament_image_read_pfm(const char *path, float *rgba, size_t pixels, uint32_t *width, uint32_t *height)
{
    if (!path || !width || !height) return ORN_ERROR_INVALID_VALUE;
    std::vector<uint8_t> b;
    if (!read_file(path, b)) return ORN_ERROR_IO;
    // three whitespace-separated header tokens after the two-byte tag
    size_t at = 0;
    auto token = [&](std::string &t) {
        t.clear();
        while (at < b.size() && std::isspace(b[at])) at++;
        while (at < b.size() && !std::isspace(b[at])) t.push_back(static_cast<char>(b[at++]));
        return !t.empty();
    };
    std::string tag, ws, hs, ss;
    if (!token(tag) || (tag != "PF" && tag != "Pf") || !token(ws) || !token(hs) || !token(ss) || at >= b.size())
        return ORN_ERROR_INVALID_FILE;
    at++;  // the single whitespace byte that ends the header
    char *e1 = nullptr, *e2 = nullptr, *e3 = nullptr;
    const unsigned long wv = std::strtoul(ws.c_str(), &e1, 10), hv = std::strtoul(hs.c_str(), &e2, 10);
    const double scale = std::strtod(ss.c_str(), &e3);
    if (*e1 || *e2 || *e3 || wv == 0 || hv == 0 || wv > 65535 || hv > 65535 || scale == 0.0)
        return ORN_ERROR_INVALID_FILE;
    const uint32_t comps = tag == "PF" ? 3u : 1u;
    const size_t px = size_t(wv) * hv;
    if (b.size() - at != px * comps * 4) return ORN_ERROR_INVALID_FILE;
    *width = static_cast<uint32_t>(wv);
    *height = static_cast<uint32_t>(hv);
    if (!rgba) return ORN_OK;
    if (pixels < px) return ORN_ERROR_SIZE_MISMATCH;
    const bool swap = scale > 0.0;  // big-endian data
    for (size_t i = 0; i < px; i++) {
        float c[3];
        for (uint32_t k = 0; k < comps; k++) {
            uint8_t q[4];
            std::memcpy(q, b.data() + at + 4 * (i * comps + k), 4);
            if (swap) {
                std::swap(q[0], q[3]);
                std::swap(q[1], q[2]);
            }
            std::memcpy(&c[k], q, 4);
        }
        rgba[4 * i] = c[0];
        rgba[4 * i + 1] = comps == 3 ? c[1] : c[0];
        rgba[4 * i + 2] = comps == 3 ? c[2] : c[0];
        rgba[4 * i + 3] = 1.0f;
    }
    return ORN_OK;
}

orn_status ornament_image_write_exr(const char *path, const float *rgba, uint32_t width, uint32_t height)
{
    if (!path || !rgba || width == 0 || height == 0 || width > 0x7fffffffu || height > 0x7fffffffu)
        return ORN_ERROR_INVALID_VALUE;
    Writer w;
    w.put(static_cast<uint32_t>(20000630));  // magic
    w.put(static_cast<uint32_t>(2));         // version 2, single-part scanline
    auto attr = [&](const char *name, const char *type, uint32_t size) {
        w.put_bytes(name, std::strlen(name) + 1);
        w.put_bytes(type, std::strlen(type) + 1);
        w.put(size);
    };
    static const char *kChannels[4] = {"A", "B", "G", "R"};  // sorted, as the format requires
    attr("channels", "chlist", 4 * (2 + 16) + 1);
    for (const char *c : kChannels) {
        w.put_bytes(c, 2);
        w.put(static_cast<int32_t>(2));  // FLOAT
        w.put(static_cast<uint32_t>(0)); // pLinear + reserved
        w.put(static_cast<int32_t>(1));  // xSampling
        w.put(static_cast<int32_t>(1));  // ySampling
    }
    w.put(static_cast<uint8_t>(0));
    attr("compression", "compression", 1);
    w.put(static_cast<uint8_t>(0));  // NO_COMPRESSION
    const int32_t box[4] = {0, 0, static_cast<int32_t>(width) - 1, static_cast<int32_t>(height) - 1};
    attr("dataWindow", "box2i", 16);
    w.put_bytes(box, 16);
    attr("displayWindow", "box2i", 16);
    w.put_bytes(box, 16);
    attr("lineOrder", "lineOrder", 1);
    w.put(static_cast<uint8_t>(0));  // INCREASING_Y
    attr("pixelAspectRatio", "float", 4);
    w.put(1.0f);
    attr("screenWindowCenter", "v2f", 8);
    w.put(0.0f);
    w.put(0.0f);
    attr("screenWindowWidth", "float", 4);
    w.put(1.0f);
    w.put(static_cast<uint8_t>(0));  // end of header
    // offset table, then one chunk per scanline: y, byte count, the
    // channels one after another; EXR row 0 is the top of the picture
    const uint64_t line_bytes = uint64_t(width) * 4 * 4;
    const uint64_t table = w.b.size();
    for (uint32_t y = 0; y < height; y++) w.put(table + 8ull * height + uint64_t(y) * (8 + line_bytes));
    for (uint32_t y = 0; y < height; y++) {
        w.put(static_cast<int32_t>(y));
        w.put(static_cast<uint32_t>(line_bytes));
        const float *row = rgba + size_t(height - 1 - y) * width * 4;
        for (int ch : {3, 2, 1, 0})  // A, B, G, R
            for (uint32_t x = 0; x < width; x++) w.put(row[4 * x + ch]);
    }
    return write_file(path, w.b) ? ORN_OK : ORN_ERROR_IO;
}

}  // extern "C"
