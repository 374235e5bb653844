// orn_nrc.hip -- neural radiance cache on gfx950: multiresolution hash-grid
// encoding, a fully fused 5 x 64 MLP on BF16 MFMA with FP32 accumulation
// (inference, and training forward + backward), Adam, and the tracer kernels
// that terminate paths at a chosen vertex (queries) and produce the
// radiance samples the cache trains on.
//
// The cache learns the reference estimator's own value of a path from a
// vertex on (ray_color, pathtracer.wgsl:95-161, with every quirk: sky
// gradient, depth exhaustion returning the throughput, metal absorption):
// a query at vertex k of a path stands for the rest of that path. Only the
// termination changes; everything before vertex k is the reference's
// arithmetic (the same device functions as the plain tracer).
//
// Network (all layers without bias; a constant-1 input feature gives the
// first layer one):
//   x0 = bf16(inputs[64])          inputs: 32 hash-grid features, direction,
//                                   normal, albedo, roughness, material type
//   x(l+1) = bf16(relu(x(l) W(l)))  l = 0..4, W(l) 64 x 64
//   y = x5 W5                       W5 64 x 16, RGB in the first 3 outputs
// Loss: relative L2, mean over samples and RGB of (y - t)^2 / (y^2 + 0.01)
// with the denominator held constant. Backward rounds each layer's output
// gradient to bf16 before its two products (FP32 accumulation), as the
// forward rounds activations.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/ornament_hip.h"
#include "orn_handles.h"
#include "orn_trace.h"

using namespace orn;
using namespace orn_rt;

namespace {

constexpr uint32_t kBlock = 256;
// internal BVH nodes a lane visits in a row before a leaf visit (bvh_hit)
constexpr uint32_t kNodeSteps = 4;
// nrc_grid_kernel: table entries per workgroup (64 KB of LDS, two feature
// sums per entry) and lanes per workgroup
constexpr uint32_t kPartEntries = 8192;
constexpr uint32_t kGridBlock = 1024;
constexpr uint32_t kIn = 64;     // network input width
constexpr uint32_t kW = 64;      // hidden width
constexpr uint32_t kOut = 16;    // output width (RGB + padding)
constexpr uint32_t kHidden = 5;  // 64 x 64 layers
constexpr uint32_t kMlpParams = kHidden * kIn * kW + kW * kOut;  // 21,504
constexpr uint32_t kOutOffset = kHidden * kIn * kW;              // W5 in the parameter vector
constexpr uint32_t kLevels = 16;
constexpr uint32_t kGridFeatures = 2 * kLevels;
constexpr uint32_t kStride = 72;  // LDS row stride in bf16 (64 + 8: rows start 4 banks apart)
constexpr uint32_t kWtElems = kHidden * kW * kStride + kOut * kStride;  // transposed, padded weights
constexpr uint32_t kTileRows = 64;                                       // samples per workgroup tile

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Hash grid layout and the scene box it covers.
struct NrcGrid {
    uint32_t log2_table;
    uint32_t res[kLevels];    // cells per axis of each level
    uint32_t dense[kLevels];  // 1: (res + 1)^3 <= table size, indexed directly
    float lo[3];
    float inv_ext[3];
};

// One cache query or training sample (80 bytes; orn_nrc_record).
struct NrcRecord {
    float pos[3];
    uint32_t pixel;
    float dir[3];
    float roughness;
    float normal[3];
    uint32_t type;
    float albedo[3];
    uint32_t flags;
    float weight[3];  // throughput of the path up to the query, or the training target
    uint32_t pad;
};
static_assert(sizeof(NrcRecord) == 80, "record layout");

// ---------------------------------------------------------------------------
// numerics shared by every kernel (and restated by the oracle)
// ---------------------------------------------------------------------------

// float -> bf16, round to nearest even (NaN stays NaN)
__device__ __forceinline__ uint16_t to_bf16(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0u;
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

// Table entry of lattice point (cx, cy, cz) at a level of `res` cells per
// axis: direct (dense) or hashed with the Instant-NGP primes.
__device__ __forceinline__ uint32_t grid_index_at(uint32_t res, uint32_t dense, uint32_t log2_table, uint32_t cx,
                                                  uint32_t cy, uint32_t cz)
{
    if (dense) {
        const uint32_t r1 = res + 1u;
        return cx + r1 * (cy + r1 * cz);
    }
    return (cx ^ (cy * 2654435761u) ^ (cz * 805459861u)) & ((1u << log2_table) - 1u);
}

// scene position -> [0, 1]^3
__device__ __forceinline__ void unit_position(const NrcGrid &g, const float pos[3], float u[3])
{
    for (int a = 0; a < 3; a++) u[a] = fminf(fmaxf((pos[a] - g.lo[a]) * g.inv_ext[a], 0.0f), 1.0f);
}

// The cell around u at a level of `res` cells per axis: its lower lattice
// corner c0 and the fractional position fr inside it.
__device__ __forceinline__ void cell_at(uint32_t res, const float u[3], uint32_t c0[3], float fr[3])
{
    const float r = static_cast<float>(res);
    for (int a = 0; a < 3; a++) {
        const float x = u[a] * r;
        const uint32_t c = min(static_cast<uint32_t>(floorf(x)), res - 1u);
        c0[a] = c;
        fr[a] = x - static_cast<float>(c);
    }
}

// Trilinear weight of corner c (bit 0 = x, bit 1 = y, bit 2 = z): (wx * wy) * wz.
__device__ __forceinline__ float corner_weight(const float fr[3], uint32_t c)
{
    const uint32_t bx = c & 1u, by = (c >> 1) & 1u, bz = (c >> 2) & 1u;
    const float wx = bx ? fr[0] : 1.0f - fr[0];
    const float wy = by ? fr[1] : 1.0f - fr[1];
    const float wz = bz ? fr[2] : 1.0f - fr[2];
    return (wx * wy) * wz;
}

// The 8 corners of the cell around u at level l, with their trilinear weights.
__device__ __forceinline__ void level_corners(const NrcGrid &g, uint32_t l, const float u[3], uint32_t idx[8],
                                              float w[8])
{
    uint32_t c0[3];
    float fr[3];
    cell_at(g.res[l], u, c0, fr);
    for (uint32_t c = 0; c < 8u; c++) {
        idx[c] = grid_index_at(g.res[l], g.dense[l], g.log2_table, c0[0] + (c & 1u), c0[1] + ((c >> 1) & 1u),
                               c0[2] + ((c >> 2) & 1u));
        w[c] = corner_weight(fr, c);
    }
}

// Word 3 of a raw 32-bit buffer descriptor on the gfx9 family.
constexpr int kRawBufferWord3 = 0x00020000;

// The encoding splits a record's 16 levels over 4 lanes: lane part (lane % 4)
// of the record's group takes levels 4 part .. 4 part + 3. Their sizes are
// fixed per lane, so a kernel loads them once (lane_levels) instead of per
// record.
struct LaneLevels {
    uint32_t res[4];
    uint32_t dense[4];
};
__device__ __forceinline__ LaneLevels lane_levels(const NrcGrid &g, uint32_t part)
{
    LaneLevels ll;
    for (uint32_t j = 0; j < 4u; j++) {
        ll.res[j] = g.res[4u * part + j];
        ll.dense[j] = g.dense[4u * part + j];
    }
    return ll;
}

// The 8 grid features of levels 4 part .. 4 part + 3 at u, into f[2 j],
// f[2 j + 1] for level 4 part + j: all 32 corner gathers are issued before
// any is used, then each level's features are summed in corner order
// (feature += w * value).
__device__ __forceinline__ void lane_features(const NrcGrid &g, const LaneLevels &ll, const __half2 *table,
                                              uint32_t part, const float u[3], float f[8])
{
    uint32_t idx[4][8];  // entry index in the whole table (levels x 2^log2_table < 2^32)
    float fr[4][3];
    for (uint32_t j = 0; j < 4u; j++) {
        uint32_t c0[3];
        cell_at(ll.res[j], u, c0, fr[j]);
        const uint32_t level_base = (4u * part + j) << g.log2_table;
        for (uint32_t c = 0; c < 8u; c++)
            idx[j][c] = level_base + grid_index_at(ll.res[j], ll.dense[j], g.log2_table, c0[0] + (c & 1u),
                                                   c0[1] + ((c >> 1) & 1u), c0[2] + ((c >> 2) & 1u));
    }
    // the gathers go through a buffer descriptor: a 32-bit offset register
    // per load instead of a 64-bit address pair
    const __amdgpu_buffer_rsrc_t trs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<__half2 *>(table), 0, static_cast<int>((kLevels << g.log2_table) * 4u), kRawBufferWord3);
    __half2 v[4][8];
    for (uint32_t j = 0; j < 4u; j++)
        for (uint32_t c = 0; c < 8u; c++)
            v[j][c] = __builtin_bit_cast(__half2, __builtin_amdgcn_raw_buffer_load_b32(trs, idx[j][c] * 4u, 0, 0));
    for (uint32_t j = 0; j < 4u; j++) {
        float f0 = 0.0f, f1 = 0.0f;
        for (uint32_t c = 0; c < 8u; c++) {
            const float2 fv = __half22float2(v[j][c]);
            const float w = corner_weight(fr[j], c);
            f0 = f0 + w * fv.x;
            f1 = f1 + w * fv.y;
        }
        f[2u * j] = f0;
        f[2u * j + 1u] = f1;
    }
}

// Network inputs 32..63 of a record: direction and normal mapped to [0, 1],
// albedo, roughness, material type one-hot, a constant 1, zeros.
__device__ __forceinline__ float extra_input(const NrcRecord &r, uint32_t k)
{
    switch (k) {
    case 32: case 33: case 34: return r.dir[k - 32] * 0.5f + 0.5f;
    case 35: case 36: case 37: return r.normal[k - 35] * 0.5f + 0.5f;
    case 38: case 39: case 40: return r.albedo[k - 38];
    case 41: return r.roughness;
    case 42: case 43: case 44: case 45: return r.type == k - 42 ? 1.0f : 0.0f;
    case 46: return 1.0f;
    default: return 0.0f;
    }
}

// ---------------------------------------------------------------------------
// MFMA tiles. v_mfma_f32_16x16x16_bf16: lane l supplies A[l % 16][4 (l / 16) + j]
// and B[4 (l / 16) + j][l % 16] (j = 0..3) and holds D[4 (l / 16) + j][l % 16].
// ---------------------------------------------------------------------------

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// gfx950's v_mfma_f32_16x16x32_bf16: twice the K of the instruction above per
// issue. Lane l supplies A[l % 16][8 (l / 16) + j] and B[8 (l / 16) + j][l % 16]
// (j = 0..7) and holds D as above.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 row8(const uint16_t *p)  // 8 consecutive (16-byte aligned)
{
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(p));
}

__device__ __forceinline__ bf16x4 row4(const uint16_t *p)  // 4 consecutive (8-byte aligned)
{
    const uint2 v = *reinterpret_cast<const uint2 *>(p);
    bf16x4 r;
    r.x = static_cast<short>(v.x & 0xffffu);
    r.y = static_cast<short>(v.x >> 16);
    r.z = static_cast<short>(v.y & 0xffffu);
    r.w = static_cast<short>(v.y >> 16);
    return r;
}

__device__ __forceinline__ bf16x4 col4(const uint16_t *p)  // 4 rows of one column
{
    bf16x4 r;
    r.x = static_cast<short>(p[0]);
    r.y = static_cast<short>(p[kStride]);
    r.z = static_cast<short>(p[2 * kStride]);
    r.w = static_cast<short>(p[3 * kStride]);
    return r;
}

// Forward of one layer for one 16-sample tile `it`: Y[i][n] = sum_k X[i][k]
// W[k][n] over n_tiles 16-wide output tiles, from the transposed weights Wt
// ([n][k] rows). With relu the result is stored as bf16 into Y; otherwise
// the FP32 tile is returned through out (n_tiles must be 1).
__device__ __forceinline__ void layer_forward(const uint16_t *Wt, uint32_t n_tiles, const uint16_t *X, uint16_t *Y,
                                              uint32_t it, uint32_t lane, f32x4 *out)
{
    const uint32_t r = lane % 16u, q = lane / 16u;
    for (uint32_t nt = 0; nt < n_tiles; nt++) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        for (uint32_t ks = 0; ks < 2u; ks++) {  // K = 64 in two 32-wide steps
            const bf16x8 a = row8(Wt + (nt * 16u + r) * kStride + ks * 32u + 8u * q);
            const bf16x8 b = row8(X + (it * 16u + r) * kStride + ks * 32u + 8u * q);
            acc = mfma32(a, b, acc);
        }
        // acc[j] = Y[it*16 + r][nt*16 + 4q + j]
        if (out) {
            *out = acc;
        } else {  // relu leaves no NaN: pack_bf16_rne == to_bf16
            uint2 v;
            v.x = pack_bf16_rne(fmaxf(acc[0], 0.0f), fmaxf(acc[1], 0.0f));
            v.y = pack_bf16_rne(fmaxf(acc[2], 0.0f), fmaxf(acc[3], 0.0f));
            *reinterpret_cast<uint2 *>(Y + (it * 16u + r) * kStride + nt * 16u + 4u * q) = v;
        }
    }
}

// Weight gradient tile (kt, nt) of one layer over the 64 samples of a tile:
// dW[k][n] = sum_i X[i][k] dY[i][n]. Lane holds dW[kt*16 + 4q + j][nt*16 + r].
__device__ __forceinline__ f32x4 weight_grad(const uint16_t *X, const uint16_t *dY, uint32_t kt, uint32_t nt,
                                             uint32_t lane)
{
    const uint32_t r = lane % 16u, q = lane / 16u;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t is = 0; is < kTileRows / 16u; is++) {
        const bf16x4 a = col4(X + (is * 16u + 4u * q) * kStride + kt * 16u + r);
        const bf16x4 b = col4(dY + (is * 16u + 4u * q) * kStride + nt * 16u + r);
        acc = mfma(a, b, acc);
    }
    return acc;
}

// Input gradient tile (it, kt): dX[i][k] = sum_n dY[i][n] W[k][n] over
// n_steps 16-wide steps of n. Lane holds dX[it*16 + 4q + j][kt*16 + r].
__device__ __forceinline__ f32x4 input_grad(const uint16_t *dY, const uint16_t *Wt, uint32_t n_steps, uint32_t it,
                                            uint32_t kt, uint32_t lane)
{
    const uint32_t r = lane % 16u, q = lane / 16u;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t ns = 0; ns < n_steps; ns++) {
        const bf16x4 a = row4(dY + (it * 16u + r) * kStride + ns * 16u + 4u * q);
        const bf16x4 b = col4(Wt + (ns * 16u + 4u * q) * kStride + kt * 16u + r);
        acc = mfma(a, b, acc);
    }
    return acc;
}

__device__ __forceinline__ const uint16_t *layer_wt(const uint16_t *wt, uint32_t l) { return wt + l * kW * kStride; }

// Record of tile row s: index base + s (modulo wrap when wrap > 0); rows at
// or past `valid` hold no sample.
__device__ __forceinline__ uint32_t tile_record(uint32_t base, uint32_t s, uint32_t wrap)
{
    return wrap ? (base + s) % wrap : base + s;
}

// 8 consecutive bf16 values to a 16-byte aligned place
__device__ __forceinline__ void store8_bf16(uint16_t *dst, const float v[8])
{
    uint32_t w[4];
    for (int k = 0; k < 4; k++) w[k] = uint32_t(to_bf16(v[2 * k])) | (uint32_t(to_bf16(v[2 * k + 1])) << 16);
    *reinterpret_cast<uint4 *>(dst) = make_uint4(w[0], w[1], w[2], w[3]);
}

// The network inputs of record rc (index i) into its bf16 row, lane part of
// the record's 4-lane group: grid features 8 part .. 8 part + 7 (levels
// 4 part .. 4 part + 3) and inputs 32 + 8 part .. 32 + 8 part + 7.
// fp32_out (optional) receives them before rounding, at row i. zero: the row
// holds no record and the lane clears its 16 entries.
__device__ __forceinline__ void encode_record(const NrcGrid &g, const LaneLevels &ll, const __half2 *table,
                                              const NrcRecord *recs, uint32_t i, bool zero, uint32_t part,
                                              uint16_t *row, float *fp32_out)
{
    if (zero) {
        uint4 *r = reinterpret_cast<uint4 *>(row + part * 16u);
        r[0] = make_uint4(0u, 0u, 0u, 0u);
        r[1] = make_uint4(0u, 0u, 0u, 0u);
        return;
    }
    const NrcRecord rc = recs[i];  // the whole record in one round trip, ahead of the gathers
    float u[3];
    unit_position(g, rc.pos, u);
    float f[8], e[8];
    lane_features(g, ll, table, part, u, f);
    for (uint32_t k = 0; k < 8u; k++) e[k] = extra_input(rc, kGridFeatures + part * 8u + k);
    store8_bf16(row + part * 8u, f);
    store8_bf16(row + kGridFeatures + part * 8u, e);
    if (fp32_out) {
        for (uint32_t k = 0; k < 8u; k++) {
            fp32_out[size_t(i) * kIn + part * 8u + k] = f[k];
            fp32_out[size_t(i) * kIn + kGridFeatures + part * 8u + k] = e[k];
        }
    }
}

// Network inputs of a 64-sample tile into X (bf16 rows); rows without a
// sample are zeros. 4 lanes per sample (encode_record).
__device__ __forceinline__ void encode_tile(const NrcGrid &g, const __half2 *table, const NrcRecord *recs,
                                            uint32_t base, uint32_t valid, uint32_t wrap, uint16_t *X,
                                            float *fp32_out)
{
    const uint32_t s = threadIdx.x / 4u, part = threadIdx.x % 4u;
    const LaneLevels ll = lane_levels(g, part);
    encode_record(g, ll, table, recs, s < valid ? tile_record(base, s, wrap) : 0u, s >= valid, part,
                  X + s * kStride, fp32_out);
}

__device__ __forceinline__ void load_weights(const uint16_t *__restrict__ wt, uint16_t *s_wt)
{
    const uint4 *src = reinterpret_cast<const uint4 *>(wt);
    uint4 *dst = reinterpret_cast<uint4 *>(s_wt);
    for (uint32_t i = threadIdx.x; i < kWtElems / 8u; i += kBlock) dst[i] = src[i];
}

// Where a render's samples go: the target's accumulation and, after the
// render's last sample (post), its framebuffer. cur is State.current_iteration
// after this sample.
struct PixelSink {
    float4 *acc;
    float4 *fb;
    uint32_t width, height;
    float cur;
    int post;
    float inv_gamma;
    uint32_t flip_y;
};

// One finished sample of pixel i: State.nextIteration's accumulation
// (pathtracer.wgsl:64-79) and, on the last sample, main_post_processing
// (pathtracer.wgsl:81-93).
__device__ __forceinline__ void add_sample(const PixelSink &o, uint32_t i, f3 c)
{
    float4 a;
    if (o.cur > 1.0f) {
        a = o.acc[i];
        a.x = a.x + c.x;
        a.y = a.y + c.y;
        a.z = a.z + c.z;
        a.w = a.w + 1.0f;
    } else {
        a = make_float4(c.x, c.y, c.z, 1.0f);
    }
    o.acc[i] = a;
    if (!o.post) return;
    float4 v;
    v.x = pow_f(fminf(fmaxf(a.x / o.cur, 0.0f), 1.0f), o.inv_gamma);
    v.y = pow_f(fminf(fmaxf(a.y / o.cur, 0.0f), 1.0f), o.inv_gamma);
    v.z = pow_f(fminf(fmaxf(a.z / o.cur, 0.0f), 1.0f), o.inv_gamma);
    v.w = fminf(fmaxf(a.w / o.cur, 0.0f), 1.0f);
    const uint32_t x = i % o.width, y = i / o.width;
    o.fb[o.flip_y < 1u ? i : o.width * (o.height - y - 1u) + x] = v;
}

__device__ __forceinline__ void wave_lds_order() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront"); }

// The forward of one layer for the 16-sample tile in rows X[0..15], in
// place: every lane first loads its sample's inputs (lane (r, q) the K
// slices 8q..8q+7 and 32+8q..32+8q+7 of row r), then the layer's four
// 16-wide output tiles are formed from the transposed weights Wt ([n][k]
// rows) and stored, relu'd and rounded to bf16, over the same row. The
// wave's loads are issued before any of its stores, so no lane reads an
// output. With out (the 16-wide output layer) the FP32 tile is returned
// instead and nothing is stored.
__device__ __forceinline__ void rows_forward(const uint16_t *Wt, uint32_t n_tiles, uint16_t *X, uint32_t lane,
                                             f32x4 *out)
{
    const uint32_t r = lane % 16u, q = lane / 16u;
    const bf16x8 b0 = row8(X + r * kStride + 8u * q);
    const bf16x8 b1 = row8(X + r * kStride + 32u + 8u * q);
    f32x4 acc[4];
    for (uint32_t nt = 0; nt < n_tiles; nt++) {
        f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
        c = mfma32(row8(Wt + (nt * 16u + r) * kStride + 8u * q), b0, c);
        c = mfma32(row8(Wt + (nt * 16u + r) * kStride + 32u + 8u * q), b1, c);
        acc[nt] = c;  // acc[nt][j] = Y[r][nt*16 + 4q + j]
    }
    if (out) {
        *out = acc[0];
        return;
    }
    for (uint32_t nt = 0; nt < n_tiles; nt++) {  // relu leaves no NaN: pack_bf16_rne == to_bf16
        uint2 v;
        v.x = pack_bf16_rne(fmaxf(acc[nt][0], 0.0f), fmaxf(acc[nt][1], 0.0f));
        v.y = pack_bf16_rne(fmaxf(acc[nt][2], 0.0f), fmaxf(acc[nt][3], 0.0f));
        *reinterpret_cast<uint2 *>(X + r * kStride + nt * 16u + 4u * q) = v;
    }
}

// ---------------------------------------------------------------------------
// Inference: predictions for n records. The workgroup stages the weights in
// LDS once; after that every wave works alone, on 16-record tiles (tile
// index blockIdx.x * kInferWaves + wave, then strides of the whole grid), in
// its own 16 LDS rows: encoding (4 lanes per record) and the six layers,
// in place, with no workgroup barrier. mode 0: raw predictions to
// out_rgb[3 i]; mode 1: the sample of the record's pixel, weight *
// max(prediction, 0), into `out` (a path that reaches its query has gathered
// no light before it: emitters end paths).
// ---------------------------------------------------------------------------
constexpr uint32_t kInferWaves = 16;
__global__ __launch_bounds__(64u * kInferWaves) void nrc_infer_kernel(NrcGrid g, const __half2 *__restrict__ table,
                                                                      const uint16_t *__restrict__ wt,
                                                                      const NrcRecord *__restrict__ recs,
                                                                      const uint32_t *__restrict__ count_ptr,
                                                                      uint32_t count_max, int mode,
                                                                      float *__restrict__ out_rgb, PixelSink out,
                                                                      float *__restrict__ inputs)
{
    __shared__ __attribute__((aligned(16))) uint16_t s_wt[kWtElems];
    __shared__ __attribute__((aligned(16))) uint16_t s_x[kInferWaves][16u * kStride];
    const uint32_t n = min(count_ptr ? *count_ptr : count_max, count_max);
    const uint32_t lane = threadIdx.x % 64u, wave = threadIdx.x / 64u;
    {
        const uint4 *src = reinterpret_cast<const uint4 *>(wt);
        uint4 *dst = reinterpret_cast<uint4 *>(s_wt);
        for (uint32_t i = threadIdx.x; i < kWtElems / 8u; i += 64u * kInferWaves) dst[i] = src[i];
    }
    __syncthreads();
    uint16_t *X = s_x[wave];
    const LaneLevels ll = lane_levels(g, lane % 4u);
    for (uint32_t tile = blockIdx.x * kInferWaves + wave; tile * 16u < n; tile += gridDim.x * kInferWaves) {
        const uint32_t first = tile * 16u;
        // encoding: lane group sr = lane / 4 holds record first + sr
        const uint32_t sr = lane / 4u;
        encode_record(g, ll, table, recs, first + sr, first + sr >= n, lane % 4u, X + sr * kStride, inputs);
        // The rows are this wave's alone: its LDS accesses complete in the
        // order it issues them, so only the compiler must keep that order (a
        // wavefront-scope fence emits no instruction).
        wave_lds_order();
        for (uint32_t l = 0; l < kHidden; l++) {
            rows_forward(layer_wt(s_wt, l), kW / 16u, X, lane, nullptr);
            wave_lds_order();
        }
        f32x4 y;
        rows_forward(layer_wt(s_wt, kHidden), 1u, X, lane, &y);
        // lanes 0..15 hold channels 0..3 of record first + lane
        const uint32_t i = first + lane;
        if (lane < 16u && i < n) {
            if (mode == 0) {
                out_rgb[3u * i] = y[0];
                out_rgb[3u * i + 1u] = y[1];
                out_rgb[3u * i + 2u] = y[2];
            } else {
                const NrcRecord &rc = recs[i];
                add_sample(out, rc.pixel,
                           mk(rc.weight[0] * fmaxf(y[0], 0.0f), rc.weight[1] * fmaxf(y[1], 0.0f),
                              rc.weight[2] * fmaxf(y[2], 0.0f)));
            }
        }
        wave_lds_order();  // the next tile's encoding overwrites the rows
    }
}

// ---------------------------------------------------------------------------
// Training: forward + backward of one 64-sample tile per workgroup over a
// batch of min(batch, record count) samples, records first, first + 1, ...
// taken modulo the record count (a frame with fewer records than a step
// needs steps over each of them once; none: an empty step; workgroups past
// the batch write nothing). Writes the tile's weight gradient to
// partials[blockIdx.x], the tile's loss to loss_sum, and for batch sample b
// its unit position (sample_u[b]), the loss gradient of its grid features
// (sample_dx[l * dx_stride + b], both features of level l) and the ranges
// of kPartEntries table entries its 8 corners at level l fall in
// (sample_mask[l * dx_stride + b]: bit (range % 64) of each; 0 when the
// gradient is 0), from which nrc_grid_kernel forms the grid gradient without
// global atomics.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void nrc_train_kernel(NrcGrid g, const __half2 *__restrict__ table,
                                                           const uint16_t *__restrict__ wt,
                                                           const NrcRecord *__restrict__ recs,
                                                           const uint32_t *__restrict__ count_ptr, uint32_t count_max,
                                                           uint32_t first, uint32_t batch,
                                                           float *__restrict__ partials, float4 *__restrict__ sample_u,
                                                           float2 *__restrict__ sample_dx,
                                                           uint64_t *__restrict__ sample_mask, uint32_t dx_stride,
                                                           float *__restrict__ loss_sum)
{
    __shared__ __attribute__((aligned(16))) uint16_t s_wt[kWtElems];
    __shared__ __attribute__((aligned(16))) uint16_t s_a[kHidden + 1u][kTileRows * kStride];
    __shared__ __attribute__((aligned(16))) uint16_t s_dy[2][kTileRows * kStride];
    __shared__ float s_dx0[kTileRows * kGridFeatures];
    const uint32_t total = min(count_ptr ? *count_ptr : count_max, count_max);
    const uint32_t n_batch = min(batch, total);  // a frame with fewer records steps over all of them
    const uint32_t tile_rel = blockIdx.x * kTileRows;
    if (tile_rel >= n_batch) return;  // no sample: nrc_mlp_kernel does not read this workgroup's partial
    const uint32_t valid = min(n_batch - tile_rel, kTileRows);
    const uint32_t tile_base = first + tile_rel;
    const uint32_t lane = threadIdx.x % 64u, wave = threadIdx.x / 64u;
    const uint32_t r = lane % 16u, q = lane / 16u;
    float *part = partials + size_t(blockIdx.x) * kMlpParams;
    load_weights(wt, s_wt);
    encode_tile(g, table, recs, tile_base, valid, total, s_a[0], nullptr);
    __syncthreads();
    for (uint32_t l = 0; l < kHidden; l++) {
        layer_forward(layer_wt(s_wt, l), kW / 16u, s_a[l], s_a[l + 1u], wave, lane, nullptr);
        __syncthreads();
    }
    // output layer and the loss gradient, row by row into s_dy[0] (16 columns)
    f32x4 y;
    layer_forward(layer_wt(s_wt, kHidden), 1u, s_a[kHidden], nullptr, wave, lane, &y);
    {
        const uint32_t s = wave * 16u + lane;
        float loss = 0.0f;
        if (lane < 16u) {
            const uint32_t i = s < valid ? tile_record(tile_base, s, total) : 0u;
            uint16_t *row = s_dy[0] + s * kStride;
            const float scale = n_batch ? 1.0f / (3.0f * static_cast<float>(n_batch)) : 0.0f;
            for (uint32_t c = 0; c < kOut; c++) {
                float gr = 0.0f;
                if (c < 3u && s < valid) {
                    const float t = recs[i].weight[c];
                    const float d = y[c] - t;
                    const float den = y[c] * y[c] + 0.01f;
                    loss = loss + (d * d / den) * scale;
                    gr = (2.0f * d / den) * scale;
                }
                row[c] = to_bf16(gr);
            }
        }
        for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off);
        if (lane == 0u && loss != 0.0f) atomicAdd(loss_sum, loss);
    }
    __syncthreads();
    // backward, output layer first; dY of layer l is in s_dy[b]
    uint32_t b = 0;
    for (int l = static_cast<int>(kHidden); l >= 0; l--) {
        const bool out_layer = l == static_cast<int>(kHidden);
        const uint32_t n_tiles = out_layer ? kOut / 16u : kW / 16u;
        const uint16_t *X = s_a[l];
        const uint16_t *dY = s_dy[b];
        const uint16_t *Wt = layer_wt(s_wt, static_cast<uint32_t>(l));
        // weight gradient: wave w takes k-tile w, all n-tiles
        for (uint32_t nt = 0; nt < n_tiles; nt++) {
            const f32x4 d = weight_grad(X, dY, wave, nt, lane);
            for (int j = 0; j < 4; j++) {
                const uint32_t k = wave * 16u + 4u * q + static_cast<uint32_t>(j), nn = nt * 16u + r;
                part[out_layer ? kOutOffset + k * kOut + nn : static_cast<uint32_t>(l) * kIn * kW + k * kW + nn] = d[j];
            }
        }
        // input gradient: wave w takes sample tile w, all 4 k-tiles
        for (uint32_t kt = 0; kt < kIn / 16u; kt++) {
            const f32x4 d = input_grad(dY, Wt, n_tiles, wave, kt, lane);
            for (int j = 0; j < 4; j++) {
                const uint32_t s = wave * 16u + 4u * q + static_cast<uint32_t>(j), k = kt * 16u + r;
                if (l > 0) {
                    const float m = from_bf16(X[s * kStride + k]) > 0.0f ? d[j] : 0.0f;  // relu of layer l-1
                    s_dy[b ^ 1u][s * kStride + k] = to_bf16(m);
                } else if (k < kGridFeatures) {
                    s_dx0[s * kGridFeatures + k] = d[j];
                }
            }
        }
        b ^= 1u;
        __syncthreads();
    }
    // per-sample inputs of the grid gradient: lane group (sample s, level
    // quarter) writes 4 levels; consecutive samples are consecutive addresses
    {
        const uint32_t s = threadIdx.x % kTileRows, part4 = threadIdx.x / kTileRows;
        if (s < valid) {
            const uint32_t b = tile_rel + s;
            float u[3];
            unit_position(g, recs[tile_record(tile_base, s, total)].pos, u);
            if (part4 == 0u) sample_u[b] = make_float4(u[0], u[1], u[2], 0.0f);
            for (uint32_t l = part4 * 4u; l < part4 * 4u + 4u; l++) {
                const float2 d = make_float2(s_dx0[s * kGridFeatures + 2u * l], s_dx0[s * kGridFeatures + 2u * l + 1u]);
                sample_dx[size_t(l) * dx_stride + b] = d;
                uint64_t m = 0;
                if (d.x != 0.0f || d.y != 0.0f) {
                    uint32_t idx[8];
                    float w[8];
                    level_corners(g, l, u, idx, w);
                    for (int c = 0; c < 8; c++) m |= 1ull << ((idx[c] / kPartEntries) & 63u);
                }
                sample_mask[size_t(l) * dx_stride + b] = m;
            }
        }
    }
}

struct AdamStep {
    float lr, beta1, beta2, eps;
    float c1, c2;  // 1 - beta1^t, 1 - beta2^t
};

__device__ __forceinline__ float adam(float p, float g, float &m, float &v, const AdamStep &a)
{
    m = a.beta1 * m + (1.0f - a.beta1) * g;
    v = a.beta2 * v + (1.0f - a.beta2) * (g * g);
    const float mh = m / a.c1, vh = v / a.c2;
    return p - a.lr * mh / (sqrtf(vh) + a.eps);
}

// What the gradient kernels do with a batch's gradient.
enum GradMode : int {
    kRefresh = 0,   // no gradient: only rewrite the working copies from the masters
    kGradient = 1,  // store the gradient (orn_nrc_gradients)
    kUpdate = 2,    // one Adam step
    kUpdateMasters = 3,  // one Adam step on the fp32 masters only; the working
                         // copies are rewritten later (refresh_params)
};

// MLP parameters: dW = sum over the training workgroups' partials
// (partials[group][j], summed in a fixed order: wave s of the workgroup adds
// groups s, s + 16, ... with four loads in flight, then the 16 wave sums are
// added in order), then store it or take an Adam step, and rewrite the
// transposed, padded bf16 copy the kernels read. One workgroup covers 64
// parameters; the parameter count is a multiple of 64.
constexpr uint32_t kMlpSlices = 16;
static_assert(kMlpParams % 64u == 0u, "MLP parameters per workgroup");
__global__ __launch_bounds__(64u * kMlpSlices) void nrc_mlp_kernel(const float *__restrict__ partials, uint32_t groups,
                                                                   const uint32_t *__restrict__ count_ptr,
                                                                   uint32_t count_max, uint32_t batch, AdamStep a,
                                                                   float *__restrict__ w, float *__restrict__ grad,
                                                                   float *__restrict__ m, float *__restrict__ v,
                                                                   uint16_t *__restrict__ wt, int mode)
{
    __shared__ float s_sum[kMlpSlices][64];
    const uint32_t lane = threadIdx.x % 64u, slice = threadIdx.x / 64u;
    const uint32_t j = blockIdx.x * 64u + lane;
    if (mode != kRefresh) {
        // the training workgroups that held samples (nrc_train_kernel's batch)
        const uint32_t total = min(count_ptr ? *count_ptr : count_max, count_max);
        groups = min(groups, (min(batch, total) + kTileRows - 1u) / kTileRows);
        float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        uint32_t gi = slice;
        for (; gi + 3u * kMlpSlices < groups; gi += 4u * kMlpSlices)
            for (uint32_t u = 0; u < 4u; u++) acc[u] = acc[u] + partials[size_t(gi + u * kMlpSlices) * kMlpParams + j];
        for (; gi < groups; gi += kMlpSlices) acc[0] = acc[0] + partials[size_t(gi) * kMlpParams + j];
        s_sum[slice][lane] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    }
    __syncthreads();
    if (slice != 0u) return;
    float p = w[j];
    if (mode != kRefresh) {
        float gj = 0.0f;
        for (uint32_t k = 0; k < kMlpSlices; k++) gj = gj + s_sum[k][lane];
        grad[j] = gj;
        if (mode == kUpdate || mode == kUpdateMasters) {
            float mj = m[j], vj = v[j];
            p = adam(p, gj, mj, vj, a);
            m[j] = mj;
            v[j] = vj;
            w[j] = p;
        }
    }
    if (mode == kUpdateMasters) return;
    uint32_t dst;
    if (j < kOutOffset) {
        const uint32_t l = j / (kIn * kW), k = (j / kW) % kIn, n = j % kW;
        dst = l * kW * kStride + n * kStride + k;
    } else {
        const uint32_t k = (j - kOutOffset) / kOut, n = (j - kOutOffset) % kOut;
        dst = kHidden * kW * kStride + n * kStride + k;
    }
    wt[dst] = to_bf16(p);
}

// fp16 working copy of the grid masters, four parameters per thread (n is a
// multiple of 4: whole tables of 2^log2_table two-feature entries)
__global__ void nrc_grid_half_kernel(const float4 *__restrict__ w, uint2 *__restrict__ half_w, uint64_t n4)
{
    const uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= n4) return;
    const float4 x = w[j];
    const __half2 lo = __halves2half2(__float2half_rn(x.x), __float2half_rn(x.y));
    const __half2 hi = __halves2half2(__float2half_rn(x.z), __float2half_rn(x.w));
    half_w[j] = make_uint2(__builtin_bit_cast(uint32_t, lo), __builtin_bit_cast(uint32_t, hi));
}

// Grid gradient and update without global atomics. A workgroup owns
// kPartEntries consecutive table entries of one level and one chunk of the
// batch, and sums into LDS the contributions w_c * dL/df of the chunk's
// samples whose cell has a corner in its range (sample b: unit position
// sample_u[b], feature gradients sample_dx[l * dx_stride + b]). Each lane
// first reads the range masks (sample_mask[l * dx_stride + b]) of 64 of the
// chunk's samples and keeps, in a register, which of them can touch the
// workgroup's range; it hashes and sums only those (about an eighth of them
// at a hashed level, nearly all at a dense one). Levels whose
// table is small next to the others (the coarse, densely indexed ones, where
// nearly every corner lands in one range) split the batch into chunks so
// that every workgroup sums about as many contributions; their chunk sums
// go to scratch and nrc_grid_finish_kernel adds them in chunk order.
// Finishing a parameter: kGradient stores the sum into grad; kUpdate takes an
// Adam step, skipping parameters whose gradient is exactly 0 (they keep value
// and moments, as in Mueller et al. 2022, sec. 4: with a table much larger
// than the batch most entries are untouched), and rewrites their fp16 copies.
// The LDS sums take a parameter's contributions in no fixed order.

struct GridParts {
    uint32_t first[kLevels + 1];  // first workgroup of each level; first[kLevels] = all workgroups
    uint32_t entries[kLevels];    // table entries a level can touch
    uint32_t chunks[kLevels];     // batch chunks of each level (workgroups = ranges x chunks)
    uint32_t scratch[kLevels];    // chunked levels: first scratch entry ([chunk][entry] float2)
    uint32_t finish[kLevels + 1]; // chunked levels: first entry index in nrc_grid_finish_kernel's range
};

struct GridUpdate {
    AdamStep a;
    float *w, *m, *v, *grad;
    __half *half_w;
    int mode;
};

__device__ __forceinline__ void finish_entry(const GridUpdate &u, size_t j, float2 gs)
{
    if (u.mode == kGradient) {
        u.grad[j] = gs.x;
        u.grad[j + 1u] = gs.y;
        return;
    }
    const float gv[2] = {gs.x, gs.y};
    for (uint32_t f = 0; f < 2u; f++) {
        if (gv[f] == 0.0f) continue;
        float mj = u.m[j + f], vj = u.v[j + f];
        const float p = adam(u.w[j + f], gv[f], mj, vj, u.a);
        u.m[j + f] = mj;
        u.v[j + f] = vj;
        u.w[j + f] = p;
        if (u.mode == kUpdate) u.half_w[j + f] = __float2half_rn(p);
    }
}

__global__ __launch_bounds__(kGridBlock) void nrc_grid_kernel(NrcGrid g, GridParts parts,
                                                              const float4 *__restrict__ sample_u,
                                                              const float2 *__restrict__ sample_dx,
                                                              const uint64_t *__restrict__ sample_mask,
                                                              uint32_t dx_stride,
                                                              const uint32_t *__restrict__ count_ptr, uint32_t count_max,
                                                              uint32_t batch, float2 *__restrict__ scratch, GridUpdate up)
{
    __shared__ float2 s_acc[kPartEntries];
    uint32_t l = 0;
    while (l + 1u < kLevels && blockIdx.x >= parts.first[l + 1u]) l++;
    const uint32_t rel = blockIdx.x - parts.first[l], chunks = parts.chunks[l];
    const uint32_t e0 = (rel / chunks) * kPartEntries, chunk = rel % chunks;
    const uint32_t span = min(parts.entries[l] - e0, kPartEntries);
    for (uint32_t i = threadIdx.x; i < kPartEntries; i += kGridBlock) s_acc[i] = make_float2(0.0f, 0.0f);
    __syncthreads();
    const uint32_t total = min(count_ptr ? *count_ptr : count_max, count_max);
    const uint32_t n_batch = min(batch, total);
    const uint32_t per = (n_batch + chunks - 1u) / chunks;
    const uint32_t b_end = min(n_batch, (chunk + 1u) * per);
    const float2 *dx = sample_dx + size_t(l) * dx_stride;
    const uint64_t *msk = sample_mask + size_t(l) * dx_stride;
    const uint64_t bit = 1ull << ((e0 / kPartEntries) & 63u);
    constexpr uint32_t kScan = 64;   // samples per lane whose masks are read before any is summed
    constexpr uint32_t kAhead = 8;   // mask loads in flight together
    for (uint32_t b0 = chunk * per + threadIdx.x; b0 < b_end; b0 += kScan * kGridBlock) {
        uint64_t hits = 0;  // bit i: sample b0 + i * kGridBlock may touch the range
#pragma unroll 1
        for (uint32_t i0 = 0; i0 < kScan; i0 += kAhead) {
            uint64_t m[kAhead];
            for (uint32_t k = 0; k < kAhead; k++) {
                const uint32_t b = b0 + (i0 + k) * kGridBlock;
                m[k] = b < b_end ? msk[b] : 0ull;
            }
            for (uint32_t k = 0; k < kAhead; k++)
                if (m[k] & bit) hits |= 1ull << (i0 + k);
        }
        while (hits != 0u) {
            const uint32_t i = static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(hits))) - 1u;
            hits &= hits - 1u;
            const uint32_t b = b0 + i * kGridBlock;
            const float2 d = dx[b];
            const float4 uu = sample_u[b];
            const float u[3] = {uu.x, uu.y, uu.z};
            uint32_t idx[8];
            float wc[8];
            level_corners(g, l, u, idx, wc);
            for (int c = 0; c < 8; c++) {
                const uint32_t e = idx[c] - e0;
                if (e < span) {
                    if (d.x != 0.0f) atomicAdd(&s_acc[e].x, wc[c] * d.x);
                    if (d.y != 0.0f) atomicAdd(&s_acc[e].y, wc[c] * d.y);
                }
            }
        }
    }
    __syncthreads();
    if (chunks > 1u) {
        float2 *dst = scratch + parts.scratch[l] + size_t(chunk) * parts.entries[l] + e0;
        for (uint32_t i = threadIdx.x; i < span; i += kGridBlock) dst[i] = s_acc[i];
        return;
    }
    const size_t base = ((size_t(l) << g.log2_table) + e0) * 2u;  // parameter index of the range's first entry
    for (uint32_t i = threadIdx.x; i < span; i += kGridBlock) finish_entry(up, base + 2u * i, s_acc[i]);
}

// The chunked levels: one entry per thread, its chunk sums added in order.
__global__ void nrc_grid_finish_kernel(NrcGrid g, GridParts parts, const float2 *__restrict__ scratch, GridUpdate up)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= parts.finish[kLevels]) return;
    uint32_t l = 0;
    while (t >= parts.finish[l + 1u]) l++;  // parts.finish[l + 1] > t for some l < kLevels
    const uint32_t e = t - parts.finish[l], chunks = parts.chunks[l], entries = parts.entries[l];
    const float2 *src = scratch + parts.scratch[l] + e;
    float2 gs = src[0];
    for (uint32_t c = 1; c < chunks; c++) {
        const float2 x = src[size_t(c) * entries];
        gs.x = gs.x + x.x;
        gs.y = gs.y + x.y;
    }
    finish_entry(up, ((size_t(l) << g.log2_table) + e) * 2u, gs);
}

__global__ void nrc_zero_kernel(float *a, uint64_t n)
{
    const uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j < n) a[j] = 0.0f;
}

// ---------------------------------------------------------------------------
// Tracer side
// ---------------------------------------------------------------------------

// pathtracer.wgsl:107-109
__device__ __forceinline__ f3 sky(const f3 &d)
{
    const f3 ud = normalize(d);
    const float tt = 0.5f * (ud.y + 1.0f);
    const float a = (1.0f - tt) * 1.0f;
    return mk(a + tt * 0.5f, a + tt * 0.7f, a + tt * 1.0f);
}

// camera.wgsl:12-23 with the (x + xi) / (W - 1) jitter of pathtracer.wgsl:64-79
__device__ __forceinline__ Ray camera_ray(const orn_constant_params &p, uint32_t x, uint32_t y, uint32_t &rng)
{
    const orn_camera &cam = p.camera;
    const float u = (static_cast<float>(x) + rng_f32(rng)) / static_cast<float>(p.width - 1u);
    const float v = (static_cast<float>(y) + rng_f32(rng)) / static_cast<float>(p.height - 1u);
    const f3 disk = random_in_unit_disk(rng);
    const f3 rd = cam.lens_radius * disk;
    const f3 offset = rd.x * mk(cam.u[0], cam.u[1], cam.u[2]) + rd.y * mk(cam.v[0], cam.v[1], cam.v[2]);
    const f3 origin = mk(cam.origin[0], cam.origin[1], cam.origin[2]);
    Ray ray;
    ray.o = origin + offset;
    ray.d = ((mk(cam.lower_left_corner[0], cam.lower_left_corner[1], cam.lower_left_corner[2]) +
              u * mk(cam.horizontal[0], cam.horizontal[1], cam.horizontal[2])) +
             v * mk(cam.vertical[0], cam.vertical[1], cam.vertical[2]) - origin) -
            offset;
    return ray;
}

// Cache inputs at a shading point
template <uint32_t F>
__device__ __forceinline__ void fill_record(const DevScene &sc, const Surface &s, const Ray &ray, NrcRecord &rc)
{
    rc.pos[0] = s.p.x;
    rc.pos[1] = s.p.y;
    rc.pos[2] = s.p.z;
    const f3 ud = normalize(ray.d);
    rc.dir[0] = -ud.x;
    rc.dir[1] = -ud.y;
    rc.dir[2] = -ud.z;
    rc.normal[0] = s.normal.x;
    rc.normal[1] = s.normal.y;
    rc.normal[2] = s.normal.z;
    rc.type = s.m.type;
    rc.roughness = s.m.type == 1u ? s.m.fuzz : 0.0f;
    const f3 alb = s.m.type == 2u ? mk(1.0f, 1.0f, 1.0f)
                                  : (s.textured ? sample_texture(sc, s.m.texture, s.tu, s.tv) : s.m.albedo);
    rc.albedo[0] = alb.x;
    rc.albedo[1] = alb.y;
    rc.albedo[2] = alb.z;
    rc.flags = 0u;
    rc.pad = 0u;
}

// The reference's bounce loop (pathtracer.wgsl:99-160) from bounce b0 with
// throughput thr: returns the sample's value.
template <uint32_t F>
__device__ __forceinline__ f3 continue_path(const DevScene &sc, const orn_constant_params &p, Ray ray, f3 thr,
                                            uint32_t b0, uint32_t &rng, uint32_t *stack)
{
    for (uint32_t b = b0; b < p.depth; b++) {
        Hit h{};
        if (!bvh_hit<F, 2, kNodeSteps>(sc, p.ray_cast_epsilon, ray, stack, kBlock, h)) return thr * sky(ray.d);
        Ray next;
        f3 w;
        if (shade<F>(sc, p.textures_count, 0u, ray, h, rng, next, w)) {
            ray = next;
            thr = thr * w;
        } else {
            return thr * w;
        }
    }
    return thr;  // depth exhausted: the throughput
}

// Slot of this lane's record among the lanes of its wave that append one:
// one atomic per wave on the shared counter.
__device__ __forceinline__ uint32_t wave_append(bool want, uint32_t *count)
{
    const uint64_t m = lanes_where(want);
    if (m == 0u) return 0u;
    const uint32_t lane = threadIdx.x % 64u;
    const int leader = __ffsll(static_cast<unsigned long long>(m)) - 1;
    uint32_t base = 0;
    if (lane == static_cast<uint32_t>(leader)) base = atomicAdd(count, static_cast<uint32_t>(__popcll(m)));
    base = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(base), leader));
    return base + static_cast<uint32_t>(__popcll(m & ((1ull << lane) - 1ull)));
}

// One sample per pixel: the reference path up to the hit at bounce k, where a
// scattering surface becomes a cache query (the sample's value is then
// throughput x prediction, added by the inference kernel); paths that end
// earlier (sky, emitter, absorption, depth) keep their exact value.
// (k = 0 has its own kernel, nrc_primary_kernel.)
template <uint32_t F>
__global__ __launch_bounds__(kBlock) void nrc_paths_kernel(DevScene sc, orn_constant_params p, uint32_t k,
                                                           uint32_t *__restrict__ rng_buf, PixelSink out,
                                                           NrcRecord *__restrict__ queries,
                                                           uint32_t *__restrict__ query_count)
{
    extern __shared__ uint32_t lds_stack[];
    uint32_t *stack = lds_stack + threadIdx.x;
    // a workgroup covers a 16 x 16 tile, a wave a 16 x 4 strip (coherent rays)
    const uint32_t tiles_x = (p.width + 15u) / 16u;
    const uint32_t x = (blockIdx.x % tiles_x) * 16u + threadIdx.x % 16u;
    const uint32_t y = (blockIdx.x / tiles_x) * 16u + threadIdx.x / 16u;
    const bool active = x < p.width && y < p.height;
    const uint32_t pix = y * p.width + x;
    bool query = false;
    NrcRecord rc;
    if (active) {
        uint32_t rng = rng_buf[pix];
        Ray ray = camera_ray(p, x, y, rng);
        f3 thr = mk(1.0f, 1.0f, 1.0f);
        f3 color = thr;
        bool done = false;
        for (uint32_t b = 0; b < p.depth && !done; b++) {
            Hit h{};
            if (!bvh_hit<F, 2, kNodeSteps>(sc, p.ray_cast_epsilon, ray, stack, kBlock, h)) {
                color = thr * sky(ray.d);
                done = true;
                break;
            }
            const Surface s = surface<F>(sc, p.textures_count, 0u, ray, h);
            if (b == k && s.m.type <= 2u) {
                fill_record<F>(sc, s, ray, rc);
                rc.pixel = pix;
                rc.weight[0] = thr.x;
                rc.weight[1] = thr.y;
                rc.weight[2] = thr.z;
                query = true;
                color = mk(0.0f, 0.0f, 0.0f);
                done = true;
                break;
            }
            Ray next;
            f3 w;
            if (scatter<F>(sc, s, ray, rng, next, w)) {
                ray = next;
                thr = thr * w;
            } else {
                color = thr * w;
                done = true;
            }
        }
        if (!done) color = thr;
        if (!query) add_sample(out, pix, color);  // else the inference kernel adds it
        rng_buf[pix] = rng;
    }
    const uint32_t slot = wave_append(query, query_count);
    if (query) queries[slot] = rc;
}

// The k = 0 case: the query is the primary hit, so each camera ray is traced
// once and nothing scatters. A workgroup covers a 16 x 16 tile, each wave an
// 8 x 8 quarter of it. In scenes without meshes the wave traces its 64
// coherent camera rays as one packet (bvh_hit_packet: every ray's own
// closest hit, bit for bit; 1.6x faster than one traversal per lane on the
// random-spheres scene). Finely tessellated meshes send the rays of a tile
// down different leaves, where a packet walks the union of their paths
// (6x slower on the instanced-icosphere scene), so there each lane
// traverses alone (LDS stack). Workgroup b takes tile tile_order[b]: the
// tiles whose slowest wave took longest in the previous launch go first, so
// that the launch does not end on a costly tile dispatched last (camera
// paths move little from frame to frame); each wave records its own time in
// tile_cost for the next launch's order.
template <uint32_t F>
__global__ __launch_bounds__(kBlock) void nrc_primary_kernel(DevScene sc, orn_constant_params p,
                                                             uint32_t *__restrict__ rng_buf, PixelSink out,
                                                             NrcRecord *__restrict__ queries,
                                                             uint32_t *__restrict__ query_count,
                                                             const uint32_t *__restrict__ tile_order,
                                                             uint32_t *__restrict__ tile_cost)
{
    constexpr bool kPacket = !(F & kHasMeshes);
    extern __shared__ uint32_t lds_stack[];
    const uint32_t t0 = device_ticks();
    const uint32_t tile = tile_order[blockIdx.x];
    const uint32_t tiles_x = (p.width + 15u) / 16u;
    if (tile >= tiles_x * ((p.height + 15u) / 16u)) return;  // (a permutation of the tiles: never)
    const uint32_t wave = threadIdx.x / 64u, lane = threadIdx.x % 64u;
    const uint32_t x = (tile % tiles_x) * 16u + (wave % 2u) * 8u + lane % 8u;
    const uint32_t y = (tile / tiles_x) * 16u + (wave / 2u) * 8u + lane / 8u;
    const bool active = x < p.width && y < p.height;
    const uint32_t pix = y * p.width + x;
    uint32_t rng = 0u;
    Ray ray;
    ray.o = mk(0.0f, 0.0f, 0.0f);
    ray.d = mk(1.0f, 1.0f, 1.0f);
    if (active) {
        rng = rng_buf[pix];
        ray = camera_ray(p, x, y, rng);
    }
    Hit h{};
    bool hit = false;
    if constexpr (kPacket) {
        hit = bvh_hit_packet<F>(sc, p.ray_cast_epsilon, ray, active && p.depth > 0u, sc.stack_slots, h);
    } else if (active && p.depth > 0u) {
        hit = bvh_hit<F, 2, kNodeSteps>(sc, p.ray_cast_epsilon, ray, lds_stack + threadIdx.x, kBlock, h);
    }
    bool query = false;
    NrcRecord rc;
    if (active) {
        const f3 thr = mk(1.0f, 1.0f, 1.0f);
        f3 color = thr;  // depth 0: the loop never runs and the throughput is returned
        if (p.depth > 0u) {
            if (!hit) {
                color = thr * sky(ray.d);
            } else {
                const Surface srf = surface<F>(sc, p.textures_count, 0u, ray, h);
                if (srf.m.type <= 2u) {
                    fill_record<F>(sc, srf, ray, rc);
                    rc.pixel = pix;
                    rc.weight[0] = 1.0f;
                    rc.weight[1] = 1.0f;
                    rc.weight[2] = 1.0f;
                    query = true;
                    color = mk(0.0f, 0.0f, 0.0f);
                } else {  // material_emit (material.wgsl:76-89)
                    color = srf.m.type == 3u
                                ? (srf.textured ? sample_texture(sc, srf.m.texture, srf.tu, srf.tv) : srf.m.albedo)
                                : mk(0.0f, 0.0f, 0.0f);
                    color = thr * color;
                }
            }
        }
        if (!query) add_sample(out, pix, color);  // else the inference kernel adds it
        rng_buf[pix] = rng;
    }
    const uint32_t slot = wave_append(query, query_count);
    if (query) queries[slot] = rc;
    if (lane == 0u)
        __hip_atomic_fetch_max(tile_cost + tile, device_ticks() - t0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The launch order of nrc_primary_kernel's tiles: a stable counting sort of
// the tiles on the quarter-octave class of their recorded cost, costliest
// first (tile order within a class; all-zero costs give the tile order),
// by one workgroup, in chunks of 1,024 tiles; the costs are cleared for the
// launch that records the next ones.
constexpr uint32_t kOrderBlock = 1024;
__global__ __launch_bounds__(kOrderBlock) void nrc_tile_order_kernel(uint32_t *__restrict__ cost, uint32_t tiles,
                                                                     uint32_t *__restrict__ order)
{
    __shared__ uint32_t hist[kCostClasses];
    __shared__ uint32_t base[kCostClasses];  // next position of each class
    __shared__ uint32_t wave_count[kOrderBlock / 64u][kCostClasses];
    const uint32_t tid = threadIdx.x, wave = tid / 64u, lane = tid % 64u;
    if (tid < kCostClasses) hist[tid] = 0u;
    __syncthreads();
    for (uint32_t i = tid; i < tiles; i += kOrderBlock) atomicAdd(&hist[cost_class(cost[i])], 1u);
    __syncthreads();
    if (tid < kCostClasses) {  // positions taken by the classes above
        uint32_t above = 0;
        for (uint32_t c = tid + 1u; c < kCostClasses; c++) above += hist[c];
        base[tid] = above;
    }
    for (uint32_t c0 = 0; c0 < tiles; c0 += kOrderBlock) {
        for (uint32_t i = tid; i < (kOrderBlock / 64u) * kCostClasses; i += kOrderBlock)
            (&wave_count[0][0])[i] = 0u;
        __syncthreads();
        const uint32_t i = c0 + tid;
        const bool valid = i < tiles;
        const uint32_t cls = valid ? cost_class(cost[i]) : kCostClasses;
        if (valid) cost[i] = 0u;
        // rank among the wave's lanes of the same class, in lane order
        uint32_t rank = 0;
        uint64_t todo = lanes_where(valid);
        while (todo != 0u) {
            const uint32_t c = static_cast<uint32_t>(
                __builtin_amdgcn_readlane(static_cast<int>(cls), __ffsll(static_cast<unsigned long long>(todo)) - 1));
            const uint64_t m = lanes_where(cls == c);
            if (cls == c) rank = static_cast<uint32_t>(__popcll(m & ((1ull << lane) - 1ull)));
            if (lane == 0u) wave_count[wave][c] = static_cast<uint32_t>(__popcll(m));
            todo &= ~m;
        }
        __syncthreads();
        if (valid) {
            for (uint32_t w = 0; w < wave; w++) rank += wave_count[w][cls];
            order[base[cls] + rank] = i;
        }
        __syncthreads();
        if (tid < kCostClasses) {
            uint32_t chunk = 0;
            for (uint32_t w = 0; w < kOrderBlock / 64u; w++) chunk += wave_count[w][tid];
            base[tid] += chunk;
        }
        __syncthreads();  // wave_count is cleared for the next chunk
    }
}

__device__ __forceinline__ uint32_t mix32(uint32_t x)  // lowbias32 (Wellons)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// Training samples: paths from random pixels with their own RNG streams;
// at the hit of bounce k on a scattering surface the inputs are recorded and
// the path continues with the reference estimator; the record's target is
// the value of the path from that vertex on (throughput 1 at the vertex).
// Path j leaves its record in stage[j] with keep[j] = 1 (else keep[j] = 0),
// and each workgroup its record count in block_keep; nrc_scan_kernel and
// nrc_compact_kernel then pack the records in path order, so a frame's
// samples do not depend on the order in which waves finish.
template <uint32_t F>
__global__ __launch_bounds__(kBlock) void nrc_train_paths_kernel(DevScene sc, orn_constant_params p, uint32_t k,
                                                                 uint32_t n_paths, uint32_t frame,
                                                                 NrcRecord *__restrict__ stage,
                                                                 uint32_t *__restrict__ keep_flags,
                                                                 uint32_t *__restrict__ block_keep)
{
    extern __shared__ uint32_t lds_stack[];
    uint32_t *stack = lds_stack + threadIdx.x;
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    bool keep = false;
    NrcRecord rc;
    if (j < n_paths) {
        const uint32_t npx = p.width * p.height;
        const uint32_t pix = mix32(j ^ mix32(frame * 0x9e3779b9u + 0x632be5abu)) % npx;
        uint32_t rng = mix32(pix + mix32(frame ^ 0x85ebca6bu) + j * 0xc2b2ae35u);
        Ray ray = camera_ray(p, pix % p.width, pix / p.width, rng);
        for (uint32_t b = 0; b <= k && b < p.depth; b++) {
            Hit h{};
            if (!bvh_hit<F, 2, kNodeSteps>(sc, p.ray_cast_epsilon, ray, stack, kBlock, h)) break;
            const Surface s = surface<F>(sc, p.textures_count, 0u, ray, h);
            Ray next;
            f3 w;
            if (b == k) {
                if (s.m.type > 2u) break;  // emitters are exact, never queried
                fill_record<F>(sc, s, ray, rc);
                rc.pixel = pix;
                f3 v;
                if (scatter<F>(sc, s, ray, rng, next, w)) v = continue_path<F>(sc, p, next, w, b + 1u, rng, stack);
                else v = w;
                rc.weight[0] = v.x;
                rc.weight[1] = v.y;
                rc.weight[2] = v.z;
                keep = true;
                break;
            }
            if (!scatter<F>(sc, s, ray, rng, next, w)) break;
            ray = next;
        }
        keep_flags[j] = keep ? 1u : 0u;
        if (keep) stage[j] = rc;
    }
    __shared__ uint32_t s_kept[kBlock / 64u];
    const uint64_t m = lanes_where(keep);
    if (threadIdx.x % 64u == 0u) s_kept[threadIdx.x / 64u] = static_cast<uint32_t>(__popcll(m));
    __syncthreads();
    if (threadIdx.x == 0u) {
        uint32_t c = 0;
        for (uint32_t w = 0; w < kBlock / 64u; w++) c += s_kept[w];
        block_keep[blockIdx.x] = c;
    }
}

// Exclusive prefix sums of the per-workgroup record counts (one workgroup of
// kScanBlock lanes, each summing a contiguous run of counts): offsets[b] is
// the first record slot of training-path workgroup b; *rec_count the total.
constexpr uint32_t kScanBlock = 1024;
__global__ __launch_bounds__(kScanBlock) void nrc_scan_kernel(const uint32_t *__restrict__ block_keep, uint32_t blocks,
                                                              uint32_t *__restrict__ offsets,
                                                              uint32_t *__restrict__ rec_count)
{
    __shared__ uint32_t s_sum[kScanBlock];
    const uint32_t per = (blocks + kScanBlock - 1u) / kScanBlock;
    const uint32_t b0 = min(threadIdx.x * per, blocks), b1 = min(b0 + per, blocks);
    uint32_t run = 0;
    for (uint32_t b = b0; b < b1; b++) run += block_keep[b];
    s_sum[threadIdx.x] = run;
    __syncthreads();
    for (uint32_t off = 1; off < kScanBlock; off <<= 1) {  // inclusive scan (Hillis-Steele)
        const uint32_t v = threadIdx.x >= off ? s_sum[threadIdx.x - off] : 0u;
        __syncthreads();
        s_sum[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t at = s_sum[threadIdx.x] - run;
    for (uint32_t b = b0; b < b1; b++) {
        offsets[b] = at;
        at += block_keep[b];
    }
    if (threadIdx.x == kScanBlock - 1u) *rec_count = s_sum[kScanBlock - 1u];
}

// Packs the kept training records in path order: the paths of
// nrc_train_paths_kernel's workgroup b start at slot offsets[b].
__global__ __launch_bounds__(kBlock) void nrc_compact_kernel(const NrcRecord *__restrict__ stage,
                                                             const uint32_t *__restrict__ keep_flags,
                                                             const uint32_t *__restrict__ offsets,
                                                             uint32_t n_paths, NrcRecord *__restrict__ recs)
{
    __shared__ uint32_t s_wave[kBlock / 64u];
    const uint32_t lane = threadIdx.x % 64u, wave = threadIdx.x / 64u;
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    const bool keep = j < n_paths && keep_flags[j] != 0u;
    const uint64_t m = lanes_where(keep);
    if (lane == 0u) s_wave[wave] = static_cast<uint32_t>(__popcll(m));
    __syncthreads();
    uint32_t in_block = 0;
    for (uint32_t w = 0; w < wave; w++) in_block += s_wave[w];
    const uint32_t slot = offsets[blockIdx.x] + in_block + static_cast<uint32_t>(__popcll(m & ((1ull << lane) - 1ull)));
    if (keep) recs[slot] = stage[j];
}

// render() bookkeeping of one sample per pixel (pathtracer.wgsl:64-79):
// current_iteration + 1, then overwrite or add.
// Clears one of the counters (queries, training records) on the stream that
// appends to it.
__global__ void nrc_reset_count_kernel(uint32_t *c)
{
    if (threadIdx.x == 0) *c = 0u;
}

}  // namespace

// ===========================================================================
// Host
// ===========================================================================

struct orn_nrc {
    orn_device_scene *scene;
    orn_nrc_config cfg;
    NrcGrid grid;
    uint64_t grid_params;  // levels x table x 2
    float *mlp, *mlp_m, *mlp_v, *mlp_grad;
    uint16_t *mlp_wt;  // working copies the kernels read
    float *grid_w, *grid_m, *grid_v, *grid_grad;
    __half *grid_half;
    // spare working copies: orn_nrc_frame writes the next ones here beside
    // the render, which reads mlp_wt / grid_half, and then swaps the pointers
    uint16_t *mlp_wt_alt;
    __half *grid_half_alt;
    float *partials;  // per training workgroup weight gradients
    uint32_t partial_groups;
    float4 *sample_u;   // per batch sample: unit position
    float2 *sample_dx;  // per level and batch sample: loss gradient of the two grid features
    uint64_t *sample_mask;  // per level and batch sample: nrc_grid_kernel ranges its corners reach
    GridParts parts;    // nrc_grid_kernel's workgroups per level
    float2 *grid_scratch;  // chunk sums of the chunked levels
    NrcRecord *train_recs;
    NrcRecord *train_stage;      // per training path (nrc_train_paths_kernel)
    uint32_t *train_keep;        // per training path: 1 if it left a record
    uint32_t *train_block_keep;  // records per training-path workgroup
    uint32_t *train_offsets;     // first record slot of each training-path workgroup
    uint32_t train_cap;
    NrcRecord *queries;
    uint32_t query_cap;  // pixels the query buffer holds
    // primary tiles queued costliest first (nrc_tile_order_kernel)
    uint32_t *tile_cost;   // per tile: GPU time of its slowest wave in the latest primary launch
    uint32_t *tile_order;  // launch order of the tiles
    uint32_t tile_cap;     // tiles the two arrays hold
    uint32_t tile_count;   // tiles of the image the costs were recorded for
    uint32_t *counters;  // [0] queries, [1] training records
    float *loss;         // [0] loss sum of the current step
    uint32_t step;       // Adam steps taken
    hipEvent_t ev[4];    // render start/stop, train start/stop
    hipEvent_t fork, join, frame_start, frame_stop;
    hipStream_t side;    // training paths of orn_nrc_frame, beside the render
    bool rendered, trained, framed;
};

namespace {

template <uint32_t F>
void launch_paths(orn_nrc *n, const DevScene &ds, const orn_constant_params &p, uint32_t *rng, const PixelSink &out,
                  hipStream_t st, size_t lds)
{
    const uint32_t tiles = blocks_for(p.width, 16u) * blocks_for(p.height, 16u);
    if (n->cfg.terminate_bounce == 0u) {
        nrc_tile_order_kernel<<<1, kOrderBlock, 0, st>>>(n->tile_cost, tiles, n->tile_order);
        nrc_primary_kernel<F><<<tiles, kBlock, (F & kHasMeshes) ? lds : 0, st>>>(ds, p, rng, out, n->queries,
                                                                                n->counters, n->tile_order,
                                                                                n->tile_cost);
    } else
        nrc_paths_kernel<F><<<tiles, kBlock, lds, st>>>(ds, p, n->cfg.terminate_bounce, rng, out, n->queries,
                                                        n->counters);
}

template <uint32_t F>
void launch_train_paths(orn_nrc *n, const DevScene &ds, const orn_constant_params &p, uint32_t frame, hipStream_t st,
                        size_t lds)
{
    const uint32_t blocks = blocks_for(n->cfg.train_paths, kBlock);
    nrc_train_paths_kernel<F><<<blocks, kBlock, lds, st>>>(ds, p, n->cfg.terminate_bounce, n->cfg.train_paths, frame,
                                                          n->train_stage, n->train_keep, n->train_block_keep);
    nrc_scan_kernel<<<1, kScanBlock, 0, st>>>(n->train_block_keep, blocks, n->train_offsets, n->counters + 1);
    nrc_compact_kernel<<<blocks, kBlock, 0, st>>>(n->train_stage, n->train_keep, n->train_offsets, n->cfg.train_paths,
                                                  n->train_recs);
}

#define ORN_NRC_DISPATCH(fn, features, ...)        \
    switch ((features) & 7u) {                     \
    case 0: fn<0>(__VA_ARGS__); break;             \
    case 1: fn<1>(__VA_ARGS__); break;             \
    case 2: fn<2>(__VA_ARGS__); break;             \
    case 3: fn<3>(__VA_ARGS__); break;             \
    case 4: fn<4>(__VA_ARGS__); break;             \
    case 5: fn<5>(__VA_ARGS__); break;             \
    case 6: fn<6>(__VA_ARGS__); break;             \
    default: fn<7>(__VA_ARGS__); break;            \
    }

void free_nrc(orn_nrc *n)
{
    (void)hipSetDevice(n->scene->ctx->device);
    (void)hipStreamSynchronize(n->scene->ctx->stream);
    void *bufs[] = {n->mlp, n->mlp_m, n->mlp_v, n->mlp_grad, n->mlp_wt, n->grid_w, n->grid_m, n->grid_v,
                    n->grid_grad, n->grid_half, n->mlp_wt_alt, n->grid_half_alt, n->partials, n->train_recs,
                    n->queries, n->counters, n->tile_cost, n->tile_order,
                    n->loss, n->sample_u, n->sample_dx, n->sample_mask, n->grid_scratch, n->train_stage, n->train_keep,
                    n->train_block_keep, n->train_offsets};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    for (hipEvent_t e : n->ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : {n->fork, n->join, n->frame_start, n->frame_stop})
        if (e) (void)hipEventDestroy(e);
    if (n->side) {
        (void)hipStreamSynchronize(n->side);
        (void)hipStreamDestroy(n->side);
    }
    orn_device_scene *s = n->scene;
    delete n;
    release_scene(s);
}

AdamStep adam_step(const orn_nrc *n)
{
    AdamStep a;
    a.lr = n->cfg.learning_rate;
    a.beta1 = n->cfg.beta1;
    a.beta2 = n->cfg.beta2;
    a.eps = n->cfg.epsilon;
    a.c1 = static_cast<float>(1.0 - std::pow(double(n->cfg.beta1), double(n->step)));
    a.c2 = static_cast<float>(1.0 - std::pow(double(n->cfg.beta2), double(n->step)));
    return a;
}

// Working copies (bf16 MLP, fp16 grid) from the masters.
// Working copies (bf16 MLP, fp16 grid) from the masters, into wt and half
// (by default the ones the kernels read).
hipError_t refresh_params(orn_nrc *n, hipStream_t st, uint16_t *wt = nullptr, __half *half = nullptr)
{
    const AdamStep a = adam_step(n);
    nrc_mlp_kernel<<<kMlpParams / 64u, 64u * kMlpSlices, 0, st>>>(nullptr, 0u, nullptr, 0u, 0u, a, n->mlp,
                                                                 n->mlp_grad, n->mlp_m, n->mlp_v,
                                                                 wt ? wt : n->mlp_wt, kRefresh);
    const uint64_t g4 = n->grid_params / 4u;  // grid_params = 2 x levels x 2^log2_table
    nrc_grid_half_kernel<<<static_cast<uint32_t>((g4 + 255) / 256), 256, 0, st>>>(
        reinterpret_cast<const float4 *>(n->grid_w), reinterpret_cast<uint2 *>(half ? half : n->grid_half), g4);
    return hipGetLastError();
}

// Forward + backward over records [first, first + batch) of buf (count from
// the device counter or count_max), then (mode) the gradients stored into
// mlp_grad and grid_grad, or an Adam step with step count n->step.
// The first half: forward + backward of the batch (reads the network,
// writes the per-workgroup partials and the per-sample grid inputs).
hipError_t train_gradients(orn_nrc *n, const NrcRecord *buf, const uint32_t *count_ptr, uint32_t count_max,
                           uint32_t first, uint32_t batch, hipStream_t st)
{
    const uint32_t groups = blocks_for(batch, kTileRows);
    if (groups > n->partial_groups) return hipErrorInvalidValue;
    nrc_train_kernel<<<groups, kBlock, 0, st>>>(n->grid, reinterpret_cast<const __half2 *>(n->grid_half), n->mlp_wt,
                                                buf, count_ptr, count_max, first, batch, n->partials, n->sample_u,
                                                n->sample_dx, n->sample_mask, n->train_cap, n->loss);
    return hipGetLastError();
}

// The second half: MLP and grid gradients from train_gradients' output,
// stored or applied (the only kernels of a step that write the network).
hipError_t apply_gradients(orn_nrc *n, const uint32_t *count_ptr, uint32_t count_max, uint32_t batch, GradMode mode,
                           hipStream_t st)
{
    const uint32_t groups = blocks_for(batch, kTileRows);
    if (groups > n->partial_groups) return hipErrorInvalidValue;
    const AdamStep a = adam_step(n);
    nrc_mlp_kernel<<<kMlpParams / 64u, 64u * kMlpSlices, 0, st>>>(n->partials, groups, count_ptr, count_max, batch,
                                                                 a, n->mlp, n->mlp_grad, n->mlp_m, n->mlp_v,
                                                                 n->mlp_wt, mode);
    GridUpdate up;
    up.a = a;
    up.w = n->grid_w;
    up.m = n->grid_m;
    up.v = n->grid_v;
    up.grad = n->grid_grad;
    up.half_w = n->grid_half;
    up.mode = mode;
    nrc_grid_kernel<<<n->parts.first[kLevels], kGridBlock, 0, st>>>(n->grid, n->parts, n->sample_u, n->sample_dx,
                                                                    n->sample_mask, n->train_cap, count_ptr,
                                                                    count_max, batch, n->grid_scratch, up);
    if (n->parts.finish[kLevels])
        nrc_grid_finish_kernel<<<blocks_for(n->parts.finish[kLevels], 256), 256, 0, st>>>(n->grid, n->parts,
                                                                                        n->grid_scratch, up);
    return hipGetLastError();
}

hipError_t gradients(orn_nrc *n, const NrcRecord *buf, const uint32_t *count_ptr, uint32_t count_max, uint32_t first,
                     uint32_t batch, GradMode mode, hipStream_t st)
{
    const hipError_t e = train_gradients(n, buf, count_ptr, count_max, first, batch, st);
    if (e != hipSuccess) return e;
    return apply_gradients(n, count_ptr, count_max, batch, mode, st);
}

// SplitMix64 for the initial parameters (reproducible per seed)
struct SplitMix {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    float uniform() { return static_cast<float>((next() >> 40) * (1.0 / 16777216.0)); }  // [0, 1)
};

}  // namespace

extern "C" {

orn_status orn_nrc_default_config(orn_nrc_config *c)
{
    if (!c) return ORN_ERROR_INVALID_VALUE;
    std::memset(c, 0, sizeof(*c));
    c->log2_table = 19;
    c->base_resolution = 16;
    c->max_resolution = 2048;
    c->terminate_bounce = 1;
    c->train_paths = 65536;
    c->train_batch = 16384;
    c->learning_rate = 1e-2f;
    c->beta1 = 0.9f;
    c->beta2 = 0.99f;
    c->epsilon = 1e-15f;
    c->seed = 1;
    return ORN_OK;
}

orn_status orn_nrc_create(orn_device_scene *scene, const orn_nrc_config *cfg, orn_nrc **out)
{
    if (!scene) return ORN_ERROR_INVALID_HANDLE;
    if (!cfg || !out) return ORN_ERROR_INVALID_VALUE;
    *out = nullptr;
    if (cfg->log2_table < 8 || cfg->log2_table > 24 || cfg->base_resolution < 2 ||
        cfg->max_resolution < cfg->base_resolution || cfg->max_resolution > (1u << 20) || cfg->train_batch == 0 ||
        cfg->train_batch % kTileRows || cfg->train_batch > (1u << 24) || cfg->train_paths > (1u << 26) ||
        !(cfg->learning_rate > 0.0f) || !(cfg->beta1 >= 0.0f && cfg->beta1 < 1.0f) ||
        !(cfg->beta2 >= 0.0f && cfg->beta2 < 1.0f) || !(cfg->epsilon >= 0.0f))
        return ORN_ERROR_INVALID_VALUE;
    ORN_TRY(hipSetDevice(scene->ctx->device));
    orn_nrc *n = new (std::nothrow) orn_nrc();
    if (!n) return ORN_ERROR_OUT_OF_MEMORY;
    n->scene = scene;
    scene->refs++;
    n->cfg = *cfg;
    NrcGrid &g = n->grid;
    g.log2_table = cfg->log2_table;
    const double growth = std::exp((std::log(double(cfg->max_resolution)) - std::log(double(cfg->base_resolution))) /
                                   double(kLevels - 1));
    const uint64_t table = 1ull << cfg->log2_table;
    for (uint32_t l = 0; l < kLevels; l++) {
        g.res[l] = std::max<uint32_t>(1u, static_cast<uint32_t>(std::floor(cfg->base_resolution * std::pow(growth, l))));
        const uint64_t r1 = g.res[l] + 1ull;
        g.dense[l] = r1 * r1 * r1 <= table ? 1u : 0u;
        n->parts.entries[l] = g.dense[l] ? static_cast<uint32_t>(r1 * r1 * r1) : static_cast<uint32_t>(table);
    }
    // nrc_grid_kernel: a level with fewer entry ranges than the largest
    // level splits the batch into as many more chunks (at most 16)
    GridParts &gp_ = n->parts;
    const uint32_t most = blocks_for(static_cast<uint32_t>(table), kPartEntries);
    uint64_t scratch_entries = 0;
    for (uint32_t l = 0; l < kLevels; l++) {
        const uint32_t ranges = blocks_for(gp_.entries[l], kPartEntries);
        gp_.chunks[l] = std::min<uint32_t>(16u, std::max<uint32_t>(1u, most / ranges));
        gp_.first[l + 1u] = gp_.first[l] + ranges * gp_.chunks[l];
        gp_.scratch[l] = static_cast<uint32_t>(scratch_entries);
        gp_.finish[l + 1u] = gp_.finish[l] + (gp_.chunks[l] > 1u ? gp_.entries[l] : 0u);
        if (gp_.chunks[l] > 1u) scratch_entries += uint64_t(gp_.chunks[l]) * gp_.entries[l];
    }
    for (int a = 0; a < 3; a++) {
        float lo = scene->bounds_lo[a], hi = scene->bounds_hi[a];
        if (!std::isfinite(lo) || !std::isfinite(hi) || !(hi > lo)) {  // a flat or empty box
            lo = std::isfinite(lo) ? lo : 0.0f;
            hi = lo + 1.0f;
        }
        g.lo[a] = lo;
        g.inv_ext[a] = 1.0f / (hi - lo);
    }
    n->grid_params = uint64_t(kLevels) * table * 2u;
    n->train_cap = std::max(cfg->train_paths, cfg->train_batch);
    n->partial_groups = blocks_for(std::max(cfg->train_batch, n->train_cap), kTileRows);
    auto fail = [&](hipError_t e) {
        free_nrc(n);
        return from_hip(e);
    };
    hipError_t e;
    const size_t gp = size_t(n->grid_params);
    if ((e = hipMalloc(&n->mlp, kMlpParams * 4)) || (e = hipMalloc(&n->mlp_m, kMlpParams * 4)) ||
        (e = hipMalloc(&n->mlp_v, kMlpParams * 4)) || (e = hipMalloc(&n->mlp_grad, kMlpParams * 4)) ||
        (e = hipMalloc(&n->mlp_wt, kWtElems * 2)) || (e = hipMalloc(&n->grid_w, gp * 4)) ||
        (e = hipMalloc(&n->grid_m, gp * 4)) || (e = hipMalloc(&n->grid_v, gp * 4)) ||
        (e = hipMalloc(&n->grid_grad, gp * 4)) || (e = hipMalloc(&n->grid_half, gp * 2)) ||
        (e = hipMalloc(&n->mlp_wt_alt, kWtElems * 2)) || (e = hipMalloc(&n->grid_half_alt, gp * 2)) ||
        (e = hipMalloc(&n->partials, size_t(n->partial_groups) * kMlpParams * 4)) ||
        (e = hipMalloc(&n->train_recs, size_t(n->train_cap) * sizeof(NrcRecord))) ||
        (e = hipMalloc(&n->train_stage, size_t(std::max(cfg->train_paths, 1u)) * sizeof(NrcRecord))) ||
        (e = hipMalloc(&n->train_keep, size_t(std::max(cfg->train_paths, 1u)) * sizeof(uint32_t))) ||
        (e = hipMalloc(&n->train_block_keep, size_t(blocks_for(std::max(cfg->train_paths, 1u), kBlock)) * 4u)) ||
        (e = hipMalloc(&n->train_offsets, size_t(blocks_for(std::max(cfg->train_paths, 1u), kBlock)) * 4u)) ||
        (e = hipMalloc(&n->sample_u, size_t(n->train_cap) * sizeof(float4))) ||
        (e = hipMalloc(&n->sample_dx, size_t(n->train_cap) * kLevels * sizeof(float2))) ||
        (e = hipMalloc(&n->sample_mask, size_t(n->train_cap) * kLevels * sizeof(uint64_t))) ||
        (scratch_entries && (e = hipMalloc(&n->grid_scratch, size_t(scratch_entries) * sizeof(float2)))) ||
        (e = hipMalloc(&n->counters, 16)) || (e = hipMalloc(&n->loss, 16)))
        return fail(e);
    for (hipEvent_t &ev : n->ev)
        if ((e = hipEventCreate(&ev))) return fail(e);
    if ((e = hipEventCreateWithFlags(&n->fork, hipEventDisableTiming)) ||
        (e = hipEventCreateWithFlags(&n->join, hipEventDisableTiming)) || (e = hipEventCreate(&n->frame_start)) ||
        (e = hipEventCreate(&n->frame_stop)) || (e = hipStreamCreateWithFlags(&n->side, hipStreamNonBlocking)))
        return fail(e);
    // initial parameters: MLP uniform in +-sqrt(6 / (fan_in + fan_out)),
    // grid uniform in +-1e-4
    SplitMix rng{cfg->seed};
    std::vector<float> w(kMlpParams), gw(gp);
    for (uint32_t j = 0; j < kMlpParams; j++) {
        const float fan = j < kOutOffset ? float(kIn + kW) : float(kW + kOut);
        w[j] = (rng.uniform() * 2.0f - 1.0f) * std::sqrt(6.0f / fan);
    }
    for (size_t j = 0; j < gp; j++) gw[j] = (rng.uniform() * 2.0f - 1.0f) * 1e-4f;
    hipStream_t st = scene->ctx->stream;
    const uint32_t zero_blocks = static_cast<uint32_t>((gp + 255) / 256);
    if ((e = hipMemcpyAsync(n->mlp, w.data(), kMlpParams * 4, hipMemcpyHostToDevice, st)) ||
        (e = hipMemcpyAsync(n->grid_w, gw.data(), gp * 4, hipMemcpyHostToDevice, st)))
        return fail(e);
    nrc_zero_kernel<<<blocks_for(kMlpParams, 256), 256, 0, st>>>(n->mlp_m, kMlpParams);
    nrc_zero_kernel<<<blocks_for(kMlpParams, 256), 256, 0, st>>>(n->mlp_v, kMlpParams);
    nrc_zero_kernel<<<zero_blocks, 256, 0, st>>>(n->grid_m, gp);
    nrc_zero_kernel<<<zero_blocks, 256, 0, st>>>(n->grid_v, gp);
    nrc_zero_kernel<<<zero_blocks, 256, 0, st>>>(n->grid_grad, gp);
    nrc_zero_kernel<<<blocks_for(kWtElems / 2u, 256), 256, 0, st>>>(reinterpret_cast<float *>(n->mlp_wt),
                                                                    kWtElems / 2u);  // row padding stays 0
    if ((e = refresh_params(n, st)) || (e = hipStreamSynchronize(st))) return fail(e);
    *out = n;
    return ORN_OK;
}

orn_status orn_nrc_destroy(orn_nrc *n)
{
    if (!n) return ORN_ERROR_INVALID_HANDLE;
    free_nrc(n);
    return ORN_OK;
}

orn_status orn_nrc_get_info(orn_nrc *n, orn_nrc_info_t *info)
{
    if (!n) return ORN_ERROR_INVALID_HANDLE;
    if (!info) return ORN_ERROR_INVALID_VALUE;
    std::memset(info, 0, sizeof(*info));
    info->levels = kLevels;
    info->log2_table = n->grid.log2_table;
    for (uint32_t l = 0; l < kLevels; l++) {
        info->resolution[l] = n->grid.res[l];
        info->dense[l] = n->grid.dense[l];
    }
    for (int a = 0; a < 3; a++) {
        info->bounds_lo[a] = n->grid.lo[a];
        info->inv_extent[a] = n->grid.inv_ext[a];
    }
    info->mlp_params = kMlpParams;
    info->grid_params = n->grid_params;
    info->adam_steps = n->step;
    return ORN_OK;
}

orn_status orn_nrc_get_params(orn_nrc *n, float *mlp, size_t mlp_len, float *grid, size_t grid_len)
{
    if (!n) return ORN_ERROR_INVALID_HANDLE;
    if ((mlp && mlp_len < kMlpParams) || (grid && grid_len < n->grid_params)) return ORN_ERROR_SIZE_MISMATCH;
    ORN_TRY(hipSetDevice(n->scene->ctx->device));
    hipStream_t st = n->scene->ctx->stream;
    if (mlp) ORN_TRY(hipMemcpyAsync(mlp, n->mlp, kMlpParams * 4, hipMemcpyDeviceToHost, st));
    if (grid) ORN_TRY(hipMemcpyAsync(grid, n->grid_w, size_t(n->grid_params) * 4, hipMemcpyDeviceToHost, st));
    ORN_TRY(hipStreamSynchronize(st));
    return ORN_OK;
}

orn_status orn_nrc_set_params(orn_nrc *n, const float *mlp, size_t mlp_len, const float *grid, size_t grid_len)
{
    if (!n) return ORN_ERROR_INVALID_HANDLE;
    if ((mlp && mlp_len != kMlpParams) || (grid && grid_len != n->grid_params)) return ORN_ERROR_SIZE_MISMATCH;
    ORN_TRY(hipSetDevice(n->scene->ctx->device));
    hipStream_t st = n->scene->ctx->stream;
    if (mlp) ORN_TRY(hipMemcpyAsync(n->mlp, mlp, kMlpParams * 4, hipMemcpyHostToDevice, st));
    if (grid) ORN_TRY(hipMemcpyAsync(n->grid_w, grid, size_t(n->grid_params) * 4, hipMemcpyHostToDevice, st));
    ORN_TRY(refresh_params(n, st));
    ORN_TRY(hipStreamSynchronize(st));
    return ORN_OK;
}

orn_status orn_nrc_query(orn_nrc *n, const orn_nrc_record *records, size_t count, float *rgb, float *inputs)
{
    if (!n) return ORN_ERROR_INVALID_HANDLE;
    if (!records || !rgb || count == 0 || count > (1u << 24)) return ORN_ERROR_INVALID_VALUE;
    ORN_TRY(hipSetDevice(n->scene->ctx->device));
    hipStream_t st = n->scene->ctx->stream;
    NrcRecord *d_rec = nullptr;
    float *d_out = nullptr, *d_in = nullptr;
    hipError_t e = hipMalloc(&d_rec, count * sizeof(NrcRecord));
    if (!e) e = hipMalloc(&d_out, count * 12);
    if (!e && inputs) e = hipMalloc(&d_in, count * kIn * 4);
    if (!e) e = hipMemcpyAsync(d_rec, records, count * sizeof(NrcRecord), hipMemcpyHostToDevice, st);
    if (!e) {
        const uint32_t tiles = blocks_for(static_cast<uint32_t>(count), kTileRows);
        nrc_infer_kernel<<<std::min<uint32_t>(blocks_for(tiles * 4u, kInferWaves), 1024u), 64u * kInferWaves, 0, st>>>(
            n->grid, reinterpret_cast<const __half2 *>(n->grid_half), n->mlp_wt, d_rec, nullptr,
            static_cast<uint32_t>(count), 0, d_out, PixelSink{}, d_in);
        e = hipGetLastError();
    }
    if (!e) e = hipMemcpyAsync(rgb, d_out, count * 12, hipMemcpyDeviceToHost, st);
    if (!e && inputs) e = hipMemcpyAsync(inputs, d_in, count * kIn * 4, hipMemcpyDeviceToHost, st);
    if (!e) e = hipStreamSynchronize(st);
    (void)hipStreamSynchronize(st);
    if (d_rec) (void)hipFree(d_rec);
    if (d_out) (void)hipFree(d_out);
    if (d_in) (void)hipFree(d_in);
    return from_hip(e);
}

// Gradients of the loss over host records (no update), or one Adam step.
static orn_status host_batch(orn_nrc *n, const orn_nrc_record *records, size_t count, int update, float *loss,
                             float *mlp_grad, float *grid_grad, size_t grid_len)
{
    if (!n) return ORN_ERROR_INVALID_HANDLE;
    if (!records || count == 0 || count > n->train_cap) return ORN_ERROR_INVALID_VALUE;
    if (grid_grad && grid_len < n->grid_params) return ORN_ERROR_SIZE_MISMATCH;
    ORN_TRY(hipSetDevice(n->scene->ctx->device));
    hipStream_t st = n->scene->ctx->stream;
    const uint32_t c = static_cast<uint32_t>(count);
    ORN_TRY(hipMemcpyAsync(n->train_recs, records, count * sizeof(NrcRecord), hipMemcpyHostToDevice, st));
    nrc_zero_kernel<<<1, 32, 0, st>>>(n->loss, 4);
    if (update) {
        n->step++;
        ORN_TRY(gradients(n, n->train_recs, nullptr, c, 0u, c, kUpdate, st));
    } else {
        // entries of dense levels past their lattice are never written by the gradient kernel
        const uint64_t gp = n->grid_params;
        nrc_zero_kernel<<<static_cast<uint32_t>((gp + 255) / 256), 256, 0, st>>>(n->grid_grad, gp);
        ORN_TRY(gradients(n, n->train_recs, nullptr, c, 0u, c, kGradient, st));
        if (mlp_grad) ORN_TRY(hipMemcpyAsync(mlp_grad, n->mlp_grad, kMlpParams * 4, hipMemcpyDeviceToHost, st));
        if (grid_grad)
            ORN_TRY(hipMemcpyAsync(grid_grad, n->grid_grad, size_t(n->grid_params) * 4, hipMemcpyDeviceToHost, st));
    }
    float l = 0.0f;
    ORN_TRY(hipMemcpyAsync(&l, n->loss, 4, hipMemcpyDeviceToHost, st));
    ORN_TRY(hipStreamSynchronize(st));
    if (loss) *loss = l;
    return ORN_OK;
}

orn_status orn_nrc_gradients(orn_nrc *n, const orn_nrc_record *records, size_t count, float *loss, float *mlp_grad,
                             float *grid_grad, size_t grid_len)
{
    return host_batch(n, records, count, 0, loss, mlp_grad, grid_grad, grid_len);
}

orn_status orn_nrc_step(orn_nrc *n, const orn_nrc_record *records, size_t count, float *loss)
{
    return host_batch(n, records, count, 1, loss, nullptr, nullptr, 0);
}

}  // extern "C"

namespace {

// One render() with cached termination, enqueued on the scene's stream.
orn_status enqueue_render(orn_nrc *n, orn_target *t, const orn_constant_params *params, uint32_t iterations)
{
    orn_device_scene *s = n->scene;
    hipStream_t st = s->ctx->stream;
    const uint32_t px = t->width * t->height;
    if (px > n->query_cap) {  // the query buffer grows with the target
        ORN_TRY(hipStreamSynchronize(st));
        if (n->queries) (void)hipFree(n->queries);
        n->queries = nullptr;
        n->query_cap = 0;
        ORN_TRY(hipMalloc(&n->queries, size_t(px) * sizeof(NrcRecord)));
        n->query_cap = px;
    }
    const uint32_t tiles = blocks_for(t->width, 16u) * blocks_for(t->height, 16u);
    if (tiles > n->tile_cap) {  // so do the tile arrays
        ORN_TRY(hipStreamSynchronize(st));
        if (n->tile_cost) (void)hipFree(n->tile_cost);
        if (n->tile_order) (void)hipFree(n->tile_order);
        n->tile_cost = n->tile_order = nullptr;
        n->tile_cap = 0;
        ORN_TRY(hipMalloc(&n->tile_cost, size_t(tiles) * sizeof(uint32_t)));
        ORN_TRY(hipMalloc(&n->tile_order, size_t(tiles) * sizeof(uint32_t)));
        n->tile_cap = tiles;
        n->tile_count = 0;
    }
    if (tiles != n->tile_count) {  // costs of another tiling order nothing
        nrc_zero_kernel<<<blocks_for(tiles, 256), 256, 0, st>>>(reinterpret_cast<float *>(n->tile_cost), tiles);
        n->tile_count = tiles;
    }
    DevScene ds = s->ds;
    ds.stack_slots = s->stack_slots;
    const size_t lds = size_t(ds.stack_slots) * kBlock * sizeof(uint32_t);
    float cur = params->current_iteration;
    const uint32_t iters = std::max(iterations, 1u);
    for (uint32_t it = 0; it < iters; it++) {
        // samples that end without a query are added by the path kernel,
        // the others by the inference kernel
        cur = cur + 1.0f;
        PixelSink out;
        out.acc = t->acc;
        out.fb = t->fb;
        out.width = t->width;
        out.height = t->height;
        out.cur = cur;
        out.post = it + 1u == iters ? 1 : 0;
        out.inv_gamma = params->inverted_gamma;
        out.flip_y = params->flip_y;
        nrc_reset_count_kernel<<<1, 64, 0, st>>>(n->counters);
        ORN_NRC_DISPATCH(launch_paths, s->features, n, ds, *params, t->rng, out, st, lds);
        const uint32_t tiles = blocks_for(px, kTileRows);
        nrc_infer_kernel<<<std::min<uint32_t>(blocks_for(tiles * 4u, kInferWaves), uint32_t(s->ctx->compute_units)),
                           64u * kInferWaves, 0, st>>>(
            n->grid, reinterpret_cast<const __half2 *>(n->grid_half), n->mlp_wt, n->queries, n->counters, px, 1,
            nullptr, out, nullptr);
    }
    ORN_TRY(hipGetLastError());
    return ORN_OK;
}

// Training paths of a frame, enqueued on stream st.
orn_status enqueue_train_paths(orn_nrc *n, const orn_constant_params *params, uint32_t frame, hipStream_t st)
{
    orn_device_scene *s = n->scene;
    DevScene ds = s->ds;
    ds.stack_slots = s->stack_slots;
    const size_t lds = size_t(ds.stack_slots) * kBlock * sizeof(uint32_t);
    nrc_reset_count_kernel<<<1, 64, 0, st>>>(n->counters + 1);
    if (n->cfg.train_paths) ORN_NRC_DISPATCH(launch_train_paths, s->features, n, ds, *params, frame, st, lds);
    ORN_TRY(hipGetLastError());
    return ORN_OK;
}

// Adam steps first..steps-1 on the frame's training records, on stream st
// (first = 0 also resets the loss sum). Step k takes records k*b, k*b + 1,
// ... of this frame's samples, modulo their count, min(b, count) of them (an
// empty frame makes a step with zero gradient). kUpdateMasters leaves the
// working copies to a later refresh_params and suits only a single step.
orn_status enqueue_steps(orn_nrc *n, uint32_t first, uint32_t steps, GradMode mode, hipStream_t st)
{
    if (first == 0u && steps > 0u) nrc_zero_kernel<<<1, 32, 0, st>>>(n->loss, 4);
    const uint32_t b = n->cfg.train_batch;
    for (uint32_t k = first; k < steps; k++) {
        n->step++;
        ORN_TRY(train_gradients(n, n->train_recs, n->counters + 1, n->train_cap, k * b, b, st));
        ORN_TRY(apply_gradients(n, n->counters + 1, n->train_cap, b, mode, st));
    }
    return ORN_OK;
}

orn_status check_target(orn_nrc *n, orn_target *t, const orn_constant_params *params)
{
    if (!n || !t) return ORN_ERROR_INVALID_HANDLE;
    if (!params) return ORN_ERROR_INVALID_VALUE;
    if (t->scene != n->scene) return ORN_ERROR_INVALID_HANDLE;
    if (params->width != t->width || params->height != t->height) return ORN_ERROR_SIZE_MISMATCH;
    return ORN_OK;
}

}  // namespace

extern "C" {

orn_status orn_nrc_render(orn_nrc *n, orn_target *t, const orn_constant_params *params, uint32_t iterations)
{
    orn_status r = check_target(n, t, params);
    if (r != ORN_OK) return r;
    ORN_TRY(hipSetDevice(n->scene->ctx->device));
    hipStream_t st = n->scene->ctx->stream;
    ORN_TRY(hipEventRecord(n->ev[0], st));
    if ((r = enqueue_render(n, t, params, iterations)) != ORN_OK) return r;
    ORN_TRY(hipEventRecord(n->ev[1], st));
    ORN_TRY(hipStreamSynchronize(st));
    n->rendered = true;
    return ORN_OK;
}

orn_status orn_nrc_train(orn_nrc *n, const orn_constant_params *params, uint32_t frame, uint32_t steps, float *loss)
{
    if (!n) return ORN_ERROR_INVALID_HANDLE;
    if (!params || params->width == 0 || params->height == 0) return ORN_ERROR_INVALID_VALUE;
    ORN_TRY(hipSetDevice(n->scene->ctx->device));
    hipStream_t st = n->scene->ctx->stream;
    ORN_TRY(hipEventRecord(n->ev[2], st));
    orn_status r = enqueue_train_paths(n, params, frame, st);
    if (r == ORN_OK) r = enqueue_steps(n, 0u, steps, kUpdate, st);
    if (r != ORN_OK) return r;
    ORN_TRY(hipEventRecord(n->ev[3], st));
    n->trained = true;
    if (loss) {
        float l = 0.0f;
        ORN_TRY(hipMemcpyAsync(&l, n->loss, 4, hipMemcpyDeviceToHost, st));
        ORN_TRY(hipStreamSynchronize(st));
        *loss = steps ? l / static_cast<float>(steps) : 0.0f;
    }
    return ORN_OK;
}

orn_status orn_nrc_frame(orn_nrc *n, orn_target *t, const orn_constant_params *params, uint32_t iterations,
                         uint32_t frame, uint32_t steps, float *loss)
{
    orn_status r = check_target(n, t, params);
    if (r != ORN_OK) return r;
    ORN_TRY(hipSetDevice(n->scene->ctx->device));
    hipStream_t st = n->scene->ctx->stream;
    ORN_TRY(hipEventRecord(n->frame_start, st));
    ORN_TRY(hipEventRecord(n->fork, st));
    // the render first, so that its long path kernel starts at once
    if ((r = enqueue_render(n, t, params, iterations)) != ORN_OK) return r;
    ORN_TRY(hipStreamWaitEvent(n->side, n->fork, 0));
    if ((r = enqueue_train_paths(n, params, frame, n->side)) != ORN_OK) return r;
    // The render reads only the working copies (bf16 MLP, fp16 grid). The
    // first step -- forward, backward and the Adam step on the fp32 masters
    // and moments -- therefore runs beside it, and so does the rewrite of the
    // spare working copies from the updated masters (what the step's own
    // kUpdate would leave in them, since each working copy is its rounded
    // master); once both streams are done the spare copies become current,
    // and any further steps follow on the scene's stream.
    if (steps > 0u) {
        if ((r = enqueue_steps(n, 0u, 1u, kUpdateMasters, n->side)) != ORN_OK) return r;
        ORN_TRY(refresh_params(n, n->side, n->mlp_wt_alt, n->grid_half_alt));
    }
    ORN_TRY(hipEventRecord(n->join, n->side));
    ORN_TRY(hipStreamWaitEvent(st, n->join, 0));
    if (steps > 0u) {
        std::swap(n->mlp_wt, n->mlp_wt_alt);
        std::swap(n->grid_half, n->grid_half_alt);
    }
    if ((r = enqueue_steps(n, 1u, steps, kUpdate, st)) != ORN_OK) return r;
    ORN_TRY(hipEventRecord(n->frame_stop, st));
    n->framed = true;
    if (loss) {
        float l = 0.0f;
        ORN_TRY(hipMemcpyAsync(&l, n->loss, 4, hipMemcpyDeviceToHost, st));
        ORN_TRY(hipStreamSynchronize(st));
        *loss = steps ? l / static_cast<float>(steps) : 0.0f;
    }
    ORN_TRY(hipStreamSynchronize(st));
    return ORN_OK;
}

orn_status orn_nrc_frame_ms(orn_nrc *n, float *ms)
{
    if (!n) return ORN_ERROR_INVALID_HANDLE;
    if (!ms) return ORN_ERROR_INVALID_VALUE;
    *ms = 0.0f;
    ORN_TRY(hipSetDevice(n->scene->ctx->device));
    if (n->framed) ORN_TRY(hipEventElapsedTime(ms, n->frame_start, n->frame_stop));
    return ORN_OK;
}

orn_status orn_nrc_last_ms(orn_nrc *n, float *render_ms, float *train_ms)
{
    if (!n) return ORN_ERROR_INVALID_HANDLE;
    ORN_TRY(hipSetDevice(n->scene->ctx->device));
    ORN_TRY(hipStreamSynchronize(n->scene->ctx->stream));
    if (render_ms) *render_ms = 0.0f;
    if (train_ms) *train_ms = 0.0f;
    if (render_ms && n->rendered) ORN_TRY(hipEventElapsedTime(render_ms, n->ev[0], n->ev[1]));
    if (train_ms && n->trained) ORN_TRY(hipEventElapsedTime(train_ms, n->ev[2], n->ev[3]));
    return ORN_OK;
}

}  // extern "C"
