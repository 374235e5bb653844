// orn_runtime.hip -- gfx950 kernels of the path-tracing hot path and the
// implementation of the C-ABI in include/ornament_hip.h.
//
// Kernel map (reference -> here):
//   main_render / main_post_processing / main (pathtracer.wgsl:23-62), looped
//   over State.iterations by the host (wgpu_backend/path_tracer.zig:252-265)
//     -> render_kernel: one launch per render() call. Each lane owns one
//        pixel and runs all of its samples back to back (the per-pixel RNG
//        chain is serial, random.wgsl:3-9), regenerating a new camera path as
//        soon as the previous one ends so that lanes of a wave stay busy
//        across samples; the framebuffer is post-processed at the end.
//   bvh_hit (bvh.wgsl:14-137) -> trace_kernel (ray queries).
//
// Launch shape: 256-lane workgroups cover 16x16 pixel tiles (a wave is a
// 16x4 strip, so its rays are coherent). Workgroup b runs on XCD b % 8; the
// tile order is remapped so that each XCD walks one contiguous band of the
// image, keeping the BVH nodes its rays touch in that XCD's own L2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <cmath>
#include <cstdlib>
#include <mutex>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "../../include/ornament_hip.h"
#include "orn_handles.h"
#include "orn_trace.h"

using namespace orn;
using namespace orn_rt;

namespace {

constexpr uint32_t kBlock = 256;
constexpr uint32_t kTile = 16;  // kTile * kTile == kBlock
constexpr uint32_t kXcds = 8;
constexpr uint32_t kMaxStack = 63;  // the reference's private stack has 64 slots
constexpr uint64_t kMaxPixels = 1ull << 28;  // per target or batch: 4 GiB of accumulation

struct KernelParams {
    orn_constant_params p;  // single image; unused by batch launches
    uint32_t iterations;
    uint32_t tiles_x;
    uint32_t tiles;
    uint32_t tiles_per_xcd;
    uint32_t regroup;       // waiting lanes that trigger a shading round
    uint32_t leaf_regroup;  // lanes at a leaf that trigger a leaf round
    uint32_t node_steps;    // internal nodes per lane between round checks
    uint32_t root;          // TLAS root reference (single image)
    uint32_t chunk;         // samples per job of a split pixel (single image)
    uint32_t chunks;        // jobs per split pixel (single image)
    uint32_t epoch;         // tag of this launch in the stage words (single image)
    uint32_t split_pct;     // percentage of each band's pixels that are split
    uint32_t pixels;        // pixels in the pixel buffers (at most 2^28)
    const uint32_t *order;  // queue position -> tile-pixel of the band, or null: tile order
    uint32_t *cost;         // per-pixel job time to record, or null
};

// One image of a batch launch. All images of a batch share one device scene
// (their arrays concatenated) and one pool of pixel buffers.
struct BatchItem {
    orn_constant_params p;  // as in a single render() of this image
    uint32_t iterations;
    uint32_t tiles_x;
    uint32_t px_base;   // first pixel of the image in the pooled buffers
    uint32_t root;      // TLAS root reference of the image's scene
    uint32_t tex_base;  // first texture of the image's scene in the shared table
    uint32_t chunk;     // samples per job of a split pixel
    uint32_t chunks;    // jobs per split pixel
    uint32_t epoch;     // tag of this launch in the stage words
};
static_assert(sizeof(BatchItem) == 176, "BatchItem layout");

// Tile work list of a batch launch: band b owns entries
// [band_begin[b], band_begin[b] + band_tiles[b]) of `tiles`, each an
// (image, tile of the image) pair; max_chunks is the largest jobs-per-pixel
// count of the images.
struct BatchTiles {
    uint32_t band_begin[kXcds];
    uint32_t band_tiles[kXcds];
    uint32_t max_chunks;
    uint32_t pad[3];
    uint2 tiles[1];  // band_begin[kXcds-1] + band_tiles[kXcds-1] entries
};

enum LaneMode : uint32_t { kTraverse = 0u, kWaiting = 1u, kDone = 2u };

// Pixel state that moves between lanes (and XCDs) within one launch is read
// and written with device-coherent accesses (they bypass the per-XCD L2s),
// so no cache flush is needed to hand a pixel from one job to the next.
__device__ __forceinline__ uint32_t load_coherent(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float load_coherent(const float *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_coherent(uint32_t *p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// The hand-off of a split pixel's state from one job to the next (possibly
// on another XCD) is ordered by hand rather than with agent-scope
// release/acquire fences, which on gfx950 also write back (buffer_wbl2) or
// invalidate (buffer_inv) the XCD's L2 on every hand-off and made the
// instanced-icosphere render (16 jobs of 2 samples per pixel) 2.2x slower.
// * Producer: the state is stored with device-scope (sc1) accesses, which go
//   through to the device's coherence point, so no L2 write-back is needed;
//   release_wait() waits until those stores have completed (vmcnt counts
//   stores on gfx9) and is a compiler barrier ("memory" clobber); only then
//   is the stage word stored (device scope).
// * Consumer: the stage word is loaded (device scope); the state is loaded
//   only after the branch on its value, with device-scope accesses that read
//   the coherence point, so no L2 invalidation is needed; acquire_order() is
//   a wavefront-scope acquire fence, which emits no instruction but keeps the
//   compiler from moving the state loads above the stage-word load.
__device__ __forceinline__ void release_wait() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }
__device__ __forceinline__ void acquire_order() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront"); }

// 16-byte accumulation entries go through a buffer descriptor, whose
// cache-policy operand (sc1: device scope) makes one 16-byte access coherent.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
constexpr int kDeviceScope = 16;                    // cache policy sc1
constexpr int kBufferWord3 = 0x00020000;            // raw 32-bit buffer, gfx9 family
__device__ __forceinline__ float4 load_coherent(__amdgpu_buffer_rsrc_t r, uint32_t i)
{
    const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, i * 16u, 0, kDeviceScope);
    return make_float4(__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z), __uint_as_float(v.w));
}
__device__ __forceinline__ void store_coherent(__amdgpu_buffer_rsrc_t r, uint32_t i, float4 a)
{
    const u32x4 v = {__float_as_uint(a.x), __float_as_uint(a.y), __float_as_uint(a.z), __float_as_uint(a.w)};
    __builtin_amdgcn_raw_buffer_store_b128(v, r, i * 16u, 0, kDeviceScope);
}

// pathtracer.wgsl:64-79 + 95-161 + 81-93, every pixel over all iterations.
//
// Jobs: the work is cut into 16x16 pixel tiles, the tiles into one band per
// XCD (workgroup b runs on XCD b % 8), and each band is a queue of jobs. The
// first pixels of a band are one job each: a lane renders all of the pixel's
// samples back to back (its RNG chain is serial, random.wgsl:3-9). The last
// split_pct percent of the band's pixels are split into jobs of `chunk`
// consecutive samples, queued all first jobs, then all second jobs, and so
// on, so that when the queue runs dry every lane is at most one short job
// from the end instead of a whole pixel. A split job starts from the
// accumulation and RNG state its predecessor left in memory; its pixel's
// stage word says how many jobs of this launch have finished (a lane that
// draws a job before its predecessor has finished waits for it). A wave
// whose band is empty moves on to the next band, so uneven bands even out.
// With a queue order (tuning cost_order), position q of a band's queue is
// the band's tile-pixel order[q] instead of the q-th: the pixels that cost
// most in the previous render come first, so the jobs drawn last are cheap
// and the lanes run out of work nearly together.
// Within a wave, a lane alternates between traversing its current ray one
// node per loop iteration and a shading round (hit shading, accumulation,
// next camera path, next job). A lane whose traversal ended waits; the wave
// runs a shading round only when at least `regroup` lanes wait (or none is
// traversing), so node visits run with most lanes busy and the long shading
// code runs for a batch of lanes at once. Each pixel's own sequence of
// operations is the reference's, so results do not depend on the schedule.
//
// BATCH: the tiles come from a work list over many images (BatchTiles), and
// each lane reads its image's parameters from the BatchItem table; the
// pixel buffers are pools indexed by px_base + pixel.
//
// POOL (tuning workgroup_pool; spheres-only scenes, BVH2, one image): before
// each run of node steps the rays at internal nodes of the workgroup's 4
// waves are pooled: every such lane writes its traversal state to its slot in
// LDS, and after a barrier the waves take the slots in order, 64 per wave, and
// run the node steps on the owners' stack columns; after a second barrier the
// owners reload node and stack top. Leaf and shading rounds stay with the
// owners, and a shading round is taken by all waves together (workgroup-wide
// trigger). Each ray's operations are unchanged, so are the results. It fills
// the node steps' lanes (lane occupancy 0.39 -> 0.50 on random spheres) but
// costs as many instructions as it saves, and the barriers cost issue
// cycles: it is an option, not the default.
// Position of the k-th (from 0) set bit of m; m has more than k set bits.
__device__ __forceinline__ uint32_t kth_set_bit(uint64_t m, uint32_t k)
{
    uint32_t pos = 0, c;
    c = __popc(static_cast<uint32_t>(m));
    if (k >= c) { k -= c; m >>= 32; pos += 32; }
    c = __popc(static_cast<uint32_t>(m) & 0xffffu);
    if (k >= c) { k -= c; m >>= 16; pos += 16; }
    c = __popc(static_cast<uint32_t>(m) & 0xffu);
    if (k >= c) { k -= c; m >>= 8; pos += 8; }
    c = __popc(static_cast<uint32_t>(m) & 0xfu);
    if (k >= c) { k -= c; m >>= 4; pos += 4; }
    c = __popc(static_cast<uint32_t>(m) & 0x3u);
    if (k >= c) { k -= c; m >>= 2; pos += 2; }
    c = static_cast<uint32_t>(m) & 1u;
    if (k >= c) pos += 1;
    return pos;
}
// node reference of a pooled traversal that has finished
constexpr uint32_t kDoneRef = 0xffffffffu;

template <uint32_t F, uint32_t W, bool BATCH, bool POOL = false>
__global__ __launch_bounds__(kBlock) void render_kernel(DevScene sc, KernelParams kp,
                                                        const BatchItem *__restrict__ items,
                                                        const BatchTiles *__restrict__ work,
                                                        float4 *__restrict__ acc_buf,
                                                        uint32_t *__restrict__ rng_buf,
                                                        float4 *__restrict__ fb,
                                                        uint32_t *__restrict__ stage,
                                                        uint32_t *__restrict__ queues)
{
    extern __shared__ uint32_t lds_stack[];
    uint32_t *stack = lds_stack + threadIdx.x;
    const uint32_t lane = threadIdx.x % 64u;
    const __amdgpu_buffer_rsrc_t acc_rsrc =
        __builtin_amdgcn_make_buffer_rsrc(acc_buf, 0, static_cast<int>(kp.pixels * 16u), kBufferWord3);

    uint32_t x = 0, y = 0;
    uint32_t pix = 0;   // pixel index in the pixel buffers
    uint32_t item = 0;  // image of the lane's pixel (BATCH)
    uint32_t job = 0;   // index of the lane's job within its pixel
    uint32_t lane_chunks = 1;  // jobs of the lane's pixel (1: a whole pixel)
    uint32_t release_pix = 0, release_tag = 0; // stage word still to be written
    uint32_t job_t0 = 0;  // device clock when the lane's job began (cost recording)
    uint32_t rng = 0;
    float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float cur = 0.0f;
    uint32_t left = 0;
    Ray ray;  // world-space ray of the current bounce
    f3 thr = mk(1.0f, 1.0f, 1.0f);
    uint32_t bounce = 0;
    Traversal tr;
    tr.addr = kLeafRef;  // a lane that is not traversing holds a leaf-flagged reference
    uint32_t mode = kWaiting;
    bool need_pixel = true;
    bool blocked = false;  // holds a job whose predecessor has not finished
    // per-image values a lane keeps while it works on a pixel (BATCH); a
    // single image takes them from the kernel arguments
    float lane_t_min = 0.0f;
    uint32_t lane_depth = 0, lane_root = 0;
    auto t_min = [&]() { return BATCH ? lane_t_min : kp.p.ray_cast_epsilon; };
    auto depth = [&]() { return BATCH ? lane_depth : kp.p.depth; };
    auto root = [&]() { return BATCH ? lane_root : kp.root; };

    auto params = [&]() -> const orn_constant_params & {
        if constexpr (BATCH) return items[item].p;
        else return kp.p;
    };

    // camera ray of a new sample (camera.wgsl:12-23): 2 jitter draws + 2 lens draws
    auto start_sample = [&]() {
        const orn_constant_params &p = params();
        const orn_camera &cam = p.camera;
        const float u = (static_cast<float>(x) + rng_f32(rng)) / static_cast<float>(p.width - 1u);
        const float v = (static_cast<float>(y) + rng_f32(rng)) / static_cast<float>(p.height - 1u);
        const f3 disk = random_in_unit_disk(rng);
        const f3 rd = cam.lens_radius * disk;
        const f3 offset = rd.x * mk(cam.u[0], cam.u[1], cam.u[2]) + rd.y * mk(cam.v[0], cam.v[1], cam.v[2]);
        const f3 origin = mk(cam.origin[0], cam.origin[1], cam.origin[2]);
        ray.o = origin + offset;
        ray.d = ((mk(cam.lower_left_corner[0], cam.lower_left_corner[1], cam.lower_left_corner[2]) +
                  u * mk(cam.horizontal[0], cam.horizontal[1], cam.horizontal[2])) +
                 v * mk(cam.vertical[0], cam.vertical[1], cam.vertical[2]) - origin) -
                offset;
        thr = mk(1.0f, 1.0f, 1.0f);
        bounce = 0;
    };
    // stage word a finished job leaves for its successor
    auto stage_tag = [&](uint32_t jobs_done) {
        return ((BATCH ? items[item].epoch : kp.epoch) << 16) | jobs_done;
    };
    auto job_ready = [&]() { return job == 0u || load_coherent(stage + pix) == stage_tag(job); };
    // End of a job: the state goes back to memory; the last job of a pixel
    // also writes the framebuffer (pathtracer.wgsl:81-93), the others
    // release the pixel to its next job.
    auto end_job = [&]() {
        if (kp.cost)  // the pixel's cost for the next render's queue order
            __hip_atomic_fetch_add(kp.cost + pix, device_ticks() - job_t0, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        if (job + 1u < lane_chunks) {  // handed on; released at the wave's next flush
            store_coherent(acc_rsrc, pix, acc);
            store_coherent(rng_buf + pix, rng);
            if (release_tag != 0u) {  // a release is still pending: write it now
                release_wait();
                store_coherent(stage + release_pix, release_tag);
            }
            release_pix = pix;
            release_tag = stage_tag(job + 1u);
            return;
        }
        acc_buf[pix] = acc;  // the pixel's last job: read next by a later launch
        rng_buf[pix] = rng;
        const orn_constant_params &p = params();
        float4 o;
        o.x = fminf(fmaxf(acc.x / cur, 0.0f), 1.0f);
        o.y = fminf(fmaxf(acc.y / cur, 0.0f), 1.0f);
        o.z = fminf(fmaxf(acc.z / cur, 0.0f), 1.0f);
        o.w = fminf(fmaxf(acc.w / cur, 0.0f), 1.0f);
        o.x = pow_f(o.x, p.inverted_gamma);
        o.y = pow_f(o.y, p.inverted_gamma);
        o.z = pow_f(o.z, p.inverted_gamma);
        const uint32_t row = p.flip_y < 1u ? y : p.height - y - 1u;
        const uint32_t base = BATCH ? items[item].px_base : 0u;
        fb[base + row * p.width + x] = o;
    };
    // Accumulate finished samples, finish samples that need no traversal
    // (depth exhausted), and set up the next traversal; a finished job asks
    // for a new one.
    auto next_traversal = [&](bool done, f3 color) {
        for (;;) {
            if (done) {
                cur = cur + 1.0f;  // State.nextIteration
                if (cur > 1.0f) {
                    acc.x = acc.x + color.x;
                    acc.y = acc.y + color.y;
                    acc.z = acc.z + color.z;
                    acc.w = acc.w + 1.0f;
                } else {
                    acc = make_float4(color.x, color.y, color.z, 1.0f);
                }
                if (--left == 0u) {
                    end_job();
                    need_pixel = true;
                    mode = kWaiting;
                    return;
                }
                start_sample();
            }
            if (bounce >= depth()) {  // depth exhausted: the throughput is returned
                color = thr;
                done = true;
                continue;
            }
            trav_begin(root(), ray, tr);
            mode = kTraverse;
            return;
        }
    };
    // Start the lane's job from the state in memory. current_iteration + the
    // job's first sample is exact (integer-valued, below 2^24: the host
    // chunks a render only then), so `cur` equals the value the reference's
    // per-sample increments reach.
    auto begin_job = [&]() {
        const orn_constant_params &p = params();
        if (kp.cost) job_t0 = device_ticks();
        if (job == 0u) {  // state left by an earlier launch
            rng = rng_buf[pix];
            acc = acc_buf[pix];
        } else {  // state handed on within this launch
            // job_ready() read the predecessor's stage word; the state loads
            // stay after that load (see acquire_order)
            acquire_order();
            rng = load_coherent(rng_buf + pix);
            acc = load_coherent(acc_rsrc, pix);
        }
        const uint32_t iters = BATCH ? items[item].iterations : kp.iterations;
        const uint32_t spp = iters > 1u ? iters : 1u;
        const uint32_t chunk = lane_chunks == 1u ? spp : (BATCH ? items[item].chunk : kp.chunk);
        const uint32_t first = job * chunk;
        cur = first == 0u ? p.current_iteration : p.current_iteration + static_cast<float>(first);
        left = min(chunk, spp - first);
        start_sample();
        next_traversal(false, thr);
    };
    // Stage words of finished jobs are written once the job's state stores
    // have completed (release_wait: one wait for the whole wave, which by the
    // next shading round is short).
    auto flush_releases = [&]() {
        if (lanes_where(release_tag != 0u) != 0u) {
            release_wait();
            if (release_tag != 0u) {
                store_coherent(stage + release_pix, release_tag);
                release_tag = 0u;
            }
        }
    };
    // jobs per band: band_px - split whole pixels, then split x max_chunks
    const uint32_t max_chunks = BATCH ? work->max_chunks : kp.chunks;
    const uint32_t split_pct = kp.split_pct;
    auto band_pixels = [&](uint32_t b) -> uint32_t {
        if constexpr (BATCH) return work->band_tiles[b] * (kTile * kTile);
        const uint32_t first = b * kp.tiles_per_xcd;
        return first >= kp.tiles ? 0u : min(kp.tiles_per_xcd, kp.tiles - first) * (kTile * kTile);
    };

    uint32_t band = blockIdx.x % kXcds;  // wave-uniform
    uint32_t bands_left = kXcds - 1u;    // bands to visit after this one
    uint32_t band_px = band_pixels(band);
    uint32_t band_split = band_px / 100u * split_pct + band_px % 100u * split_pct / 100u;
    uint32_t band_jobs = band_px - band_split + band_split * max_chunks;
    const uint32_t regroup = kp.regroup;
    const uint32_t leaf_regroup = kp.leaf_regroup;
    const uint32_t node_steps = kp.node_steps;
    bool shade_all = false;  // POOL: this iteration's workgroup-wide shading decision
    for (;;) {
        if constexpr (POOL) {
            static_assert(!(F & kHasMeshes) && W == 2u && !BATCH, "pooled traversal: spheres, BVH2, one image");
            __shared__ float p_f[7][kBlock];  // inv xyz, ox xyz, t_max of each slot
            __shared__ uint32_t p_addr[kBlock], p_top[kBlock];
            __shared__ uint64_t p_mask[kBlock / 64u];
            __shared__ uint32_t p_alive[kBlock / 64u], p_wait[kBlock / 64u], p_trav[kBlock / 64u];
            const uint32_t wave = threadIdx.x / 64u;
            const bool inner = !(tr.addr & kLeafRef);
            const uint64_t im = lanes_where(inner);
            const bool wave_alive = lanes_where(mode != kDone) != 0u;
            const uint32_t nw = static_cast<uint32_t>(__popcll(lanes_where(mode == kWaiting)));
            const uint32_t nt = static_cast<uint32_t>(__popcll(lanes_where(mode == kTraverse)));
            if (inner) {
                p_f[0][threadIdx.x] = tr.inv.x;
                p_f[1][threadIdx.x] = tr.inv.y;
                p_f[2][threadIdx.x] = tr.inv.z;
                p_f[3][threadIdx.x] = tr.ox.x;
                p_f[4][threadIdx.x] = tr.ox.y;
                p_f[5][threadIdx.x] = tr.ox.z;
                p_f[6][threadIdx.x] = tr.t_max;
                p_addr[threadIdx.x] = tr.addr;
                p_top[threadIdx.x] = tr.top;
            }
            if (lane == 0u) {
                p_mask[wave] = im;
                p_alive[wave] = wave_alive ? 1u : 0u;
                p_wait[wave] = nw;
                p_trav[wave] = nt;
            }
            __syncthreads();
            // every wave leaves together once none has work
            if ((p_alive[0] | p_alive[1] | p_alive[2] | p_alive[3]) == 0u) break;
            const uint32_t ws = p_wait[0] + p_wait[1] + p_wait[2] + p_wait[3];
            const uint32_t ts = p_trav[0] + p_trav[1] + p_trav[2] + p_trav[3];
            shade_all = ws >= 4u * regroup || (ts == 0u ? ws > 0u : 2u * ts < ws);
            const uint32_t n0 = static_cast<uint32_t>(__popcll(p_mask[0]));
            const uint32_t n1 = static_cast<uint32_t>(__popcll(p_mask[1]));
            const uint32_t n2 = static_cast<uint32_t>(__popcll(p_mask[2]));
            const uint32_t n3 = static_cast<uint32_t>(__popcll(p_mask[3]));
            const uint32_t total = n0 + n1 + n2 + n3;
            if (wave * 64u < total) {  // wave-uniform
                const uint32_t e = wave * 64u + lane;
                const bool mine = e < total;
                uint32_t slot = 0;
                Traversal wt;
                wt.addr = kDoneRef;
                wt.top = 0;
                wt.in_tlas = true;
                wt.hit = false;
                if (mine) {
                    const uint32_t p1 = n0, p2 = n0 + n1, p3 = n0 + n1 + n2;
                    const uint32_t ow = (e >= p1 ? 1u : 0u) + (e >= p2 ? 1u : 0u) + (e >= p3 ? 1u : 0u);
                    const uint32_t before = ow == 0u ? 0u : (ow == 1u ? p1 : (ow == 2u ? p2 : p3));
                    slot = ow * 64u + kth_set_bit(p_mask[ow], e - before);
                    wt.inv = mk(p_f[0][slot], p_f[1][slot], p_f[2][slot]);
                    wt.ox = mk(p_f[3][slot], p_f[4][slot], p_f[5][slot]);
                    wt.t_max = p_f[6][slot];
                    wt.addr = p_addr[slot];
                    wt.top = p_top[slot];
                }
                for (uint32_t k = 0; k < node_steps; k++) {
                    if (!(wt.addr & kLeafRef) && !trav_internal<F, W>(sc, t_min(), ray, lds_stack + slot, kBlock, wt))
                        wt.addr = kDoneRef;
                }
                if (mine) {
                    p_addr[slot] = wt.addr;
                    p_top[slot] = wt.top;
                }
            }
            __syncthreads();
            if (inner) {
                tr.addr = p_addr[threadIdx.x];
                tr.top = p_top[threadIdx.x];
                if (tr.addr == kDoneRef) {
                    mode = kWaiting;
                    tr.addr = kLeafRef;
                }
            }
        } else {
            // up to `node_steps` internal nodes per lane every iteration;
            // leaves (sphere, triangle, mesh entry: the long code) in rounds
            // of at least `leaf_regroup` lanes. An unflagged tr.addr implies
            // mode == kTraverse: lanes that stop traversing take kLeafRef.
            for (uint32_t k = 0; k < node_steps; k++) {
                if (!(tr.addr & kLeafRef) && !trav_internal<F, W>(sc, t_min(), ray, stack, kBlock, tr)) {
                    mode = kWaiting;
                    tr.addr = kLeafRef;
                }
            }
        }
        // masks of single comparisons, combined on the scalar unit
        const uint64_t tracing = lanes_where(mode == kTraverse);
        const uint64_t at_inner = lanes_where(!(tr.addr & kLeafRef));
        const uint64_t at_leaf = tracing & ~at_inner;
        if (at_leaf != 0u && (static_cast<uint32_t>(__popcll(at_leaf)) >= leaf_regroup || at_inner == 0u)) {
            if (mode == kTraverse && (tr.addr & kLeafRef) &&
                !trav_leaf<F, W>(sc, t_min(), ray, stack, kBlock, tr)) {
                mode = kWaiting;
                tr.addr = kLeafRef;
            }
        }
        const uint64_t waiting = lanes_where(mode == kWaiting);
        const uint64_t traversing = lanes_where(mode == kTraverse);
        if (waiting == 0u && traversing == 0u) {
            if constexpr (POOL) {
                // the workgroup leaves together, at the loop's top; a release
                // this wave still holds is written now, since a lane of
                // another wave of the workgroup may be waiting for it (and
                // the workgroup leaves only once that lane is done)
                flush_releases();
                continue;
            } else {
                break;
            }
        }
        // a shading round once `regroup` lanes wait, or once fewer than half
        // as many lanes still traverse as wait (the wave would otherwise step
        // a few lanes while most of it idles)
        const uint32_t n_waiting = static_cast<uint32_t>(__popcll(waiting));
        if constexpr (POOL) {
            if (!shade_all) continue;
        } else if (n_waiting < regroup && traversing != 0u &&
                   2u * static_cast<uint32_t>(__popcll(traversing)) >= n_waiting) {
            continue;
        }
        // shading round
        flush_releases();
        if (mode == kWaiting && !need_pixel) {
            if (blocked) {
                if (job_ready()) {
                    blocked = false;
                    begin_job();
                }
            } else {
                bounce++;
                // each lane's outcome first, then one next_traversal for all
                // of them (inlined once, so lanes that missed and lanes that
                // scattered share its code)
                bool done;
                f3 color;
                if (!tr.hit) {
                    const f3 ud = normalize(ray.d);
                    const float tt = 0.5f * (ud.y + 1.0f);
                    const float a = (1.0f - tt) * 1.0f;
                    done = true;
                    color = thr * mk(a + tt * 0.5f, a + tt * 0.7f, a + tt * 1.0f);
                } else {
                    Ray next;
                    f3 w;
                    const uint32_t tex_count = params().textures_count;
                    const uint32_t tex_base = BATCH ? items[item].tex_base : 0u;
                    done = !shade<F>(sc, tex_count, tex_base, ray, tr.h, rng, next, w);
                    if (done) {
                        color = thr * w;
                    } else {
                        ray = next;
                        thr = thr * w;
                        color = thr;
                    }
                }
                next_traversal(done, color);
            }
        }
        // lanes without a job take the next ones of the wave's band, one
        // atomic per wave; pixels of partial edge tiles (and, in a batch,
        // jobs past an image's last) are skipped; a band that runs out
        // sends the wave on to the next one
        for (;;) {
            const uint64_t needy = lanes_where(mode == kWaiting && need_pixel);
            if (needy == 0u) break;
            const int leader = __ffsll(static_cast<unsigned long long>(needy)) - 1;
            const uint32_t count = static_cast<uint32_t>(__popcll(needy));
            uint32_t base = 0;
            if (lane == static_cast<uint32_t>(leader)) base = atomicAdd(queues + band, count);
            base = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(base), leader));
            if (mode == kWaiting && need_pixel) {
                const uint32_t q = base + static_cast<uint32_t>(__popcll(needy & ((1ull << lane) - 1ull)));
                if (q >= band_jobs) {
                    if (bands_left == 0u) mode = kDone;  // every band has run out
                } else {
                    // whole pixel, or job c of split pixel qp
                    const uint32_t bulk = band_px - band_split;
                    uint32_t c = 0, qp = q;
                    if (q >= bulk) {
                        c = (q - bulk) / band_split;
                        qp = bulk + (q - bulk - c * band_split);
                    }
                    // position qp of the band's queue -> tile-pixel of the band
                    const uint32_t qpx =
                        kp.order ? kp.order[(BATCH ? work->band_begin[band] : band * kp.tiles_per_xcd) *
                                                (kTile * kTile) + qp]
                                 : qp;
                    const uint32_t within = qpx % (kTile * kTile);
                    uint32_t tile, tiles_x, w, h, chunks;
                    if constexpr (BATCH) {
                        const uint2 e = work->tiles[work->band_begin[band] + qpx / (kTile * kTile)];
                        item = e.x;
                        tile = e.y;
                        tiles_x = items[item].tiles_x;
                        w = items[item].p.width;
                        h = items[item].p.height;
                        chunks = items[item].chunks;
                    } else {
                        tile = band * kp.tiles_per_xcd + qpx / (kTile * kTile);
                        tiles_x = kp.tiles_x;
                        w = kp.p.width;
                        h = kp.p.height;
                        chunks = kp.chunks;
                    }
                    x = (tile % tiles_x) * kTile + within % kTile;
                    y = (tile / tiles_x) * kTile + within / kTile;
                    const bool split = qp >= bulk;
                    if (x < w && y < h && (!split || c < chunks)) {
                        pix = y * w + x;
                        lane_chunks = split ? chunks : 1u;
                        if constexpr (BATCH) {
                            const orn_constant_params &p = params();
                            pix += items[item].px_base;
                            lane_t_min = p.ray_cast_epsilon;
                            lane_depth = p.depth;
                            lane_root = items[item].root;
                        }
                        job = c;
                        need_pixel = false;
                        if (job_ready()) {
                            begin_job();
                        } else {
                            blocked = true;
                        }
                    }
                }
            }
            if (base + count > band_jobs && bands_left != 0u) {  // the band ran out (wave-uniform)
                band = band + 1u == kXcds ? 0u : band + 1u;
                bands_left--;
                band_px = band_pixels(band);
                band_split = band_px / 100u * split_pct + band_px % 100u * split_pct / 100u;
                band_jobs = band_px - band_split + band_split * max_chunks;
            }
        }
    }
    flush_releases();  // no successor may wait on a release this wave holds
}

__global__ void reset_queues_kernel(uint32_t *q)
{
    if (threadIdx.x < kXcds) q[threadIdx.x] = 0u;
}

// Queue order by cost (tuning cost_order): a stable counting sort of each
// band's tile-pixels on their cost class, highest class first. One workgroup
// per tile counts its pixels per class (order_count_kernel); one wave per
// band and class turns the counts into prefixes over the band's tiles
// (order_scan_kernel); one workgroup per tile places its pixels
// (order_place_kernel), which also clears the costs for the render that
// records the next ones. Pixels outside an image (partial edge tiles) take
// class 0 and go last; the render skips them as before. With every cost zero
// (nothing recorded yet) the order is the tile order. cost_class (the
// quarter-octave classes) is in orn_trace.h.

// The tiles a launch queues: one image's tiles in row-major order, cut into
// kXcds bands of tiles_per_xcd (work == nullptr), or a batch's work list.
struct OrderShape {
    const BatchTiles *work;
    const BatchItem *items;
    uint32_t tiles_x, tiles, tiles_per_xcd, width, height;
};

// first tile and tile count of band b
__device__ __forceinline__ uint2 band_range(const OrderShape &o, uint32_t b)
{
    if (o.work) return make_uint2(o.work->band_begin[b], o.work->band_tiles[b]);
    const uint32_t first = b * o.tiles_per_xcd;
    return make_uint2(first, first >= o.tiles ? 0u : min(o.tiles_per_xcd, o.tiles - first));
}

// band holding tile t
__device__ __forceinline__ uint32_t band_of(const OrderShape &o, uint32_t t)
{
    if (!o.work) return t / o.tiles_per_xcd;
    uint32_t b = 0;
    while (b + 1u < kXcds && t >= o.work->band_begin[b + 1u]) b++;
    return b;
}

// cost class of pixel i (0..255) of tile t, and its index in the pixel
// buffers (0xffffffff outside the image: class 0)
__device__ __forceinline__ uint32_t tile_pixel_class(const uint32_t *cost, const OrderShape &o, uint32_t t,
                                                     uint32_t i, uint32_t *pixel)
{
    uint32_t tile = t, tiles_x = o.tiles_x, w = o.width, h = o.height, base = 0;
    if (o.work) {
        const uint2 e = o.work->tiles[t];
        const BatchItem &it = o.items[e.x];
        tile = e.y;
        tiles_x = it.tiles_x;
        w = it.p.width;
        h = it.p.height;
        base = it.px_base;
    }
    const uint32_t x = (tile % tiles_x) * kTile + i % kTile;
    const uint32_t y = (tile / tiles_x) * kTile + i / kTile;
    if (x >= w || y >= h) {
        *pixel = 0xffffffffu;
        return 0u;
    }
    *pixel = base + y * w + x;
    return cost_class(cost[*pixel]);
}

__global__ __launch_bounds__(kBlock) void order_count_kernel(const uint32_t *__restrict__ cost, OrderShape o,
                                                             uint32_t *__restrict__ counts)
{
    __shared__ uint32_t hist[kCostClasses];
    if (threadIdx.x < kCostClasses) hist[threadIdx.x] = 0u;
    __syncthreads();
    uint32_t pixel;
    atomicAdd(&hist[tile_pixel_class(cost, o, blockIdx.x, threadIdx.x, &pixel)], 1u);
    __syncthreads();
    if (threadIdx.x < kCostClasses) counts[blockIdx.x * kCostClasses + threadIdx.x] = hist[threadIdx.x];
}

// grid: kXcds * kCostClasses waves; wave w scans class w % kCostClasses over
// the tiles of band w / kCostClasses, in place (exclusive), and stores the
// band's total of the class.
__global__ __launch_bounds__(kBlock) void order_scan_kernel(uint32_t *__restrict__ counts, OrderShape o,
                                                            uint32_t *__restrict__ totals)
{
    const uint32_t w = (blockIdx.x * kBlock + threadIdx.x) / 64u;
    const uint32_t lane = threadIdx.x % 64u;
    const uint32_t band = w / kCostClasses, cls = w % kCostClasses;
    const uint2 r = band_range(o, band);
    const uint32_t end = r.x + r.y;
    uint32_t carry = 0;
    for (uint32_t t0 = r.x; t0 < end; t0 += 64u) {
        const uint32_t t = t0 + lane;
        const uint32_t v = t < end ? counts[t * kCostClasses + cls] : 0u;
        uint32_t incl = v;
        for (uint32_t d = 1; d < 64u; d <<= 1) {
            const uint32_t up = __shfl_up(incl, d);
            if (lane >= d) incl += up;
        }
        if (t < end) counts[t * kCostClasses + cls] = carry + incl - v;
        carry += __shfl(incl, 63);
    }
    if (lane == 0u) totals[band * kCostClasses + cls] = carry;
}

// order[first tile of the band * 256 + position] = the band's tile-pixel
// (tile within the band * 256 + pixel within the tile) queued there
__global__ __launch_bounds__(kBlock) void order_place_kernel(uint32_t *__restrict__ cost, OrderShape o,
                                                             const uint32_t *__restrict__ prefix,
                                                             const uint32_t *__restrict__ totals,
                                                             uint32_t *__restrict__ order)
{
    __shared__ uint32_t base[kCostClasses];  // band positions taken by the classes above
    __shared__ uint32_t wave_count[kBlock / 64u][kCostClasses];
    const uint32_t tile = blockIdx.x, band = band_of(o, tile);
    const uint32_t first = band_range(o, band).x;
    const uint32_t wave = threadIdx.x / 64u, lane = threadIdx.x % 64u;
    for (uint32_t i = threadIdx.x; i < (kBlock / 64u) * kCostClasses; i += kBlock) (&wave_count[0][0])[i] = 0u;
    if (threadIdx.x < kCostClasses) {
        // exclusive scan over the classes from the highest down (two waves)
        const uint32_t c = kCostClasses - 1u - threadIdx.x;
        const uint32_t v = totals[band * kCostClasses + c];
        uint32_t incl = v;
        for (uint32_t d = 1; d < 64u; d <<= 1) {
            const uint32_t up = __shfl_up(incl, d);
            if (lane >= d) incl += up;
        }
        base[c] = incl - v;
    }
    __syncthreads();
    if (threadIdx.x >= 64u && threadIdx.x < kCostClasses)  // the second wave adds the first one's total
        base[kCostClasses - 1u - threadIdx.x] += base[kCostClasses - 64u] + totals[band * kCostClasses + kCostClasses - 64u];
    uint32_t pixel;
    const uint32_t cls = tile_pixel_class(cost, o, tile, threadIdx.x, &pixel);
    if (pixel != 0xffffffffu) cost[pixel] = 0u;
    // rank among the wave's lanes of the same class, in lane order
    uint32_t rank = 0;
    uint64_t todo = lanes_where(true);
    while (todo != 0u) {
        const uint32_t c = static_cast<uint32_t>(
            __builtin_amdgcn_readlane(static_cast<int>(cls), __ffsll(static_cast<unsigned long long>(todo)) - 1));
        const uint64_t m = lanes_where(cls == c);
        if (cls == c) rank = static_cast<uint32_t>(__popcll(m & ((1ull << lane) - 1ull)));
        if (lane == 0u) wave_count[wave][c] = static_cast<uint32_t>(__popcll(m));
        todo &= ~m;
    }
    __syncthreads();
    for (uint32_t w = 0; w < wave; w++) rank += wave_count[w][cls];
    order[first * (kTile * kTile) + base[cls] + prefix[tile * kCostClasses + cls] + rank] =
        (tile - first) * (kTile * kTile) + threadIdx.x;
}

// The three kernels over `tiles` tiles on `st`.
void enqueue_order_kernels(const OrderShape &o, uint32_t tiles, uint32_t *cost, uint32_t *counts, uint32_t *totals,
                           uint32_t *order, hipStream_t st)
{
    order_count_kernel<<<tiles, kBlock, 0, st>>>(cost, o, counts);
    order_scan_kernel<<<kXcds * kCostClasses * 64u / kBlock, kBlock, 0, st>>>(counts, o, totals);
    order_place_kernel<<<tiles, kBlock, 0, st>>>(cost, o, counts, totals, order);
}

template <uint32_t W>
__global__ __launch_bounds__(kBlock) void trace_kernel(DevScene sc, float t_min,
                                                       const float *__restrict__ rays, uint32_t n,
                                                       orn_hit *__restrict__ out)
{
    extern __shared__ uint32_t lds_stack[];
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    Ray r;
    r.o = mk(rays[6u * i], rays[6u * i + 1u], rays[6u * i + 2u]);
    r.d = mk(rays[6u * i + 3u], rays[6u * i + 4u], rays[6u * i + 5u]);
    Hit h{};
    const bool hit = bvh_hit<kHasSpheres | kHasMeshes, W>(sc, t_min, r, lds_stack + threadIdx.x, kBlock, h);
    orn_hit o{};
    o.hit = hit ? 1u : 0u;
    if (hit) {
        o.t = h.t;
        o.material_index = h.material;
        o.node_type = h.node_type;
        o.inverted_transform_id = h.inv_tid;
        if (h.node_type == 2u) {  // triangle id and barycentrics exist for mesh hits only
            o.tri_id = h.tri_id;
            o.uv[0] = h.u;
            o.uv[1] = h.v;
        }
    }
    out[i] = o;
}

// The same query with the rays of a wave traced together (bvh_hit_packet);
// every lane takes part, lanes past n without a ray.
__global__ __launch_bounds__(kBlock) void trace_packet_kernel(DevScene sc, float t_min,
                                                              const float *__restrict__ rays, uint32_t n,
                                                              orn_hit *__restrict__ out)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const bool valid = i < n;
    Ray r;
    r.o = mk(0.0f, 0.0f, 0.0f);
    r.d = mk(1.0f, 1.0f, 1.0f);
    if (valid) {
        r.o = mk(rays[6u * i], rays[6u * i + 1u], rays[6u * i + 2u]);
        r.d = mk(rays[6u * i + 3u], rays[6u * i + 4u], rays[6u * i + 5u]);
    }
    Hit h{};
    const bool hit = bvh_hit_packet<kHasSpheres | kHasMeshes>(sc, t_min, r, valid, sc.stack_slots, h);
    if (!valid) return;
    orn_hit o{};
    o.hit = hit ? 1u : 0u;
    if (hit) {
        o.t = h.t;
        o.material_index = h.material;
        o.node_type = h.node_type;
        o.inverted_transform_id = h.inv_tid;
        if (h.node_type == 2u) {  // triangle id and barycentrics exist for mesh hits only
            o.tri_id = h.tri_id;
            o.uv[0] = h.u;
            o.uv[1] = h.v;
        }
    }
    out[i] = o;
}

__global__ void math_kernel(int op, const float *a, const float *b, float *out, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float r;
    switch (op) {
    case 0: { float s, c; sin_cos(a[i], s, c); r = s; break; }
    case 1: { float s, c; sin_cos(a[i], s, c); r = c; break; }
    case 2: r = pow_f(a[i], b[i]); break;
    case 3: r = acos_f(a[i]); break;
    case 5: r = __uint_as_float(pack_bf16_rne(a[i], 0.0f) << 16); break;
    default: r = atan2_f(a[i], b[i]); break;
    }
    out[i] = r;
}

__global__ void copy_kernel(const float4 *__restrict__ src, float4 *__restrict__ dst, uint64_t n)
{
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

__global__ void fill_u32_kernel(uint32_t *a, uint32_t n, uint32_t v)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void seed_rng_kernel(uint32_t *rng, uint32_t n, uint32_t base)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) rng[i] = i + base;
}

__global__ void zero_kernel(float4 *a, float4 *b, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        a[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        b[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

}  // namespace

// ===========================================================================
// Host runtime
// ===========================================================================

namespace {

orn_status set_device(orn_ctx *ctx) { return from_hip(hipSetDevice(ctx->device)); }

template <typename T>
orn_status upload(orn_device_scene *s, const T *src, size_t n, const T **dst)
{
    const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);  // empty arrays padded to one element
    void *p = nullptr;
    ORN_TRY(hipMalloc(&p, bytes));
    s->allocations.push_back(p);
    if (n) {
        ORN_TRY(hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice));
    } else {
        ORN_TRY(hipMemset(p, 0, bytes));
    }
    *dst = static_cast<const T *>(p);
    return ORN_OK;
}

// Depth (edges on the longest root-to-leaf path) of every node. Children
// precede their parent in both arrays (bvh.zig appends children first), which
// the validation enforces, so one pass in index order suffices.
std::vector<uint32_t> node_depths(const orn_bvh_node *nodes, size_t n)
{
    std::vector<uint32_t> d(n, 0);
    for (size_t i = 0; i < n; i++)
        if (nodes[i].node_type == 0u)
            d[i] = 1u + std::max(d[nodes[i].left_or_custom_id], d[nodes[i].right_or_material_index]);
    return d;
}

// Derived BVH4: every BVH2 internal node N becomes a BVH4 node whose slots
// hold N's grandchildren [LL, LR, RL, RR] (a child that is a leaf stands for
// itself), with the boxes the BVH2 stores for them. Leaves stay in the BVH2
// arrays and are referenced as kLeafRef | index.
struct Bvh4Node {
    float lo[3][4];  // [axis][slot]
    float hi[3][4];
    uint32_t child[4];
    uint32_t pad[4];
};
static_assert(sizeof(Bvh4Node) == 128, "BVH4 node is 128 bytes");

struct Bvh4Builder {
    const orn_bvh_node *nodes;
    std::vector<Bvh4Node> out;
    std::vector<uint32_t> memo;  // BVH2 index -> reference, 0xffffffff = not built

    explicit Bvh4Builder(const orn_bvh_node *n, size_t len) : nodes(n), memo(len, 0xffffffffu) {}

    uint32_t build(uint32_t idx)
    {
        if (nodes[idx].node_type != 0u) return kLeafRef | idx;
        if (memo[idx] != 0xffffffffu) return memo[idx];
        struct Entry {
            uint32_t idx;
            const float *lo, *hi;
        } e[4];
        int n = 0;
        const orn_bvh_node &nd = nodes[idx];
        const uint32_t kids[2] = {nd.left_or_custom_id, nd.right_or_material_index};
        const float *klo[2] = {nd.lmin_or_v0, nd.rmin_or_v2}, *khi[2] = {nd.lmax_or_v1, nd.rmax_or_v3};
        for (int k = 0; k < 2; k++) {
            const orn_bvh_node &c = nodes[kids[k]];
            if (c.node_type == 0u) {
                e[n++] = {c.left_or_custom_id, c.lmin_or_v0, c.lmax_or_v1};
                e[n++] = {c.right_or_material_index, c.rmin_or_v2, c.rmax_or_v3};
            } else {
                e[n++] = {kids[k], klo[k], khi[k]};
            }
        }
        const uint32_t me = static_cast<uint32_t>(out.size());
        out.emplace_back();
        memo[idx] = me;
        Bvh4Node node{};
        for (int i = 0; i < 4; i++) {
            for (int a = 0; a < 3; a++) {
                node.lo[a][i] = i < n ? e[i].lo[a] : 0.0f;
                node.hi[a][i] = i < n ? e[i].hi[a] : 0.0f;
            }
            node.child[i] = kEmptyRef;
        }
        for (int i = 0; i < n; i++) node.child[i] = build(e[i].idx);
        out[me] = node;
        return me;
    }
};

// Worst-case pending stack entries of the traversal from a node on, for
// BVH2 (take the right child, push the left) and BVH4 (push all hit slots
// but the last). A mesh leaf of the TLAS pushes the BLAS marker and
// continues into its BLAS.
uint32_t pending2(const orn_bvh_node *n, uint32_t i, const orn_bvh_node *blas, std::vector<uint32_t> &memo_blas,
                  bool tlas)
{
    const orn_bvh_node &nd = n[i];
    if (nd.node_type == 0u) {
        const uint32_t l = pending2(n, nd.left_or_custom_id, blas, memo_blas, tlas);
        const uint32_t r = pending2(n, nd.right_or_material_index, blas, memo_blas, tlas);
        return std::max(1u + r, l);
    }
    if (tlas && nd.node_type == 2u) {
        uint32_t &m = memo_blas[nd.left_or_custom_id];
        if (m == 0xffffffffu) m = pending2(blas, nd.left_or_custom_id, blas, memo_blas, false);
        return 1u + m;
    }
    return 0u;
}

uint32_t pending4(const std::vector<Bvh4Node> &n4, uint32_t ref, const orn_bvh_node *leaves,
                  const std::vector<Bvh4Node> *blas4, std::vector<uint32_t> *memo_blas, const uint32_t *blas_ref)
{
    if (ref & kLeafRef) {
        const orn_bvh_node &nd = leaves[ref & ~kLeafRef];
        if (blas4 && nd.node_type == 2u) {
            uint32_t &m = (*memo_blas)[nd.left_or_custom_id];
            if (m == 0xffffffffu) m = pending4(*blas4, blas_ref[nd.left_or_custom_id], nullptr, nullptr, nullptr, nullptr);
            return 1u + m;
        }
        return 0u;
    }
    const Bvh4Node &nd = n4[ref];
    uint32_t worst = 0;
    for (uint32_t i = 0; i < 4; i++)
        if (nd.child[i] != kEmptyRef)
            worst = std::max(worst, i + pending4(n4, nd.child[i], leaves, blas4, memo_blas, blas_ref));
    return worst;
}

orn_status validate(const orn_scene_desc *d, uint32_t *stack_slots)
{
    if (!d->tlas_nodes || d->tlas_nodes_len == 0) return ORN_ERROR_EMPTY_SCENE;
    if (d->tlas_nodes_len >= 0x7ffffff0u || d->blas_nodes_len >= 0x7ffffff0u) return ORN_ERROR_INVALID_SCENE;
    auto need = [](const void *p, size_t n) { return n == 0 || p != nullptr; };
    if (!need(d->blas_nodes, d->blas_nodes_len) || !need(d->normals, d->normals_len) ||
        !need(d->normal_indices, d->normal_indices_len) || !need(d->uvs, d->uvs_len) ||
        !need(d->uv_indices, d->uv_indices_len) || !need(d->transforms, d->transforms_len) ||
        !need(d->materials, d->materials_len) || !need(d->textures, d->textures_len))
        return ORN_ERROR_INVALID_VALUE;
    for (size_t i = 0; i < d->normal_indices_len; i++)
        if (d->normal_indices[i] >= d->normals_len) return ORN_ERROR_INVALID_SCENE;
    for (size_t i = 0; i < d->uv_indices_len; i++)
        if (d->uv_indices[i] >= d->uvs_len) return ORN_ERROR_INVALID_SCENE;
    auto leaf_ok = [&](const orn_bvh_node &nd) {
        const uint64_t tid = nd.transform_id;
        return 2u * tid + 1u < d->transforms_len && nd.right_or_material_index < d->materials_len;
    };
    std::vector<uint8_t> blas_root_used(d->blas_nodes_len, 0);
    for (size_t i = 0; i < d->tlas_nodes_len; i++) {
        const orn_bvh_node &nd = d->tlas_nodes[i];
        switch (nd.node_type) {
        case 0u:
            if (nd.left_or_custom_id >= i || nd.right_or_material_index >= i) return ORN_ERROR_INVALID_SCENE;
            break;
        case 1u:
            if (!leaf_ok(nd)) return ORN_ERROR_INVALID_SCENE;
            break;
        case 2u:
            if (!leaf_ok(nd) || nd.left_or_custom_id >= d->blas_nodes_len) return ORN_ERROR_INVALID_SCENE;
            blas_root_used[nd.left_or_custom_id] = 1;
            break;
        default:
            return ORN_ERROR_INVALID_SCENE;  // a triangle or unknown node cannot sit in the TLAS
        }
    }
    for (size_t i = 0; i < d->blas_nodes_len; i++) {
        const orn_bvh_node &nd = d->blas_nodes[i];
        if (nd.node_type == 0u) {
            if (nd.left_or_custom_id >= i || nd.right_or_material_index >= i) return ORN_ERROR_INVALID_SCENE;
        } else if (nd.node_type == 3u) {
            const uint64_t base = 3ull * nd.left_or_custom_id;
            if (base + 2u >= d->normal_indices_len || base + 2u >= d->uv_indices_len)
                return ORN_ERROR_INVALID_SCENE;
        } else {
            return ORN_ERROR_INVALID_SCENE;  // no nested instancing
        }
    }
    for (size_t i = 0; i < d->textures_len; i++) {
        const orn_texture_desc &t = d->textures[i];
        const bool comps_ok = t.num_components == 1 || t.num_components == 2 || t.num_components == 4;
        const bool fmt_ok = t.is_hdr ? t.bytes_per_component == 2 : t.bytes_per_component == 1;
        if (!comps_ok || !fmt_ok || t.width == 0 || t.height == 0 || t.width > ORN_MAX_TEXTURE_DIMENSION ||
            t.height > ORN_MAX_TEXTURE_DIMENSION)
            return ORN_ERROR_UNSUPPORTED_TEXTURE;
        if (!t.data) return ORN_ERROR_INVALID_VALUE;
    }
    // trees deeper than the reference's 64-entry stack are rejected before
    // anything recurses over them
    const std::vector<uint32_t> td = node_depths(d->tlas_nodes, d->tlas_nodes_len);
    const std::vector<uint32_t> bd = node_depths(d->blas_nodes, d->blas_nodes_len);
    uint32_t blas_depth = 0;
    for (size_t i = 0; i < d->blas_nodes_len; i++)
        if (blas_root_used[i]) blas_depth = std::max(blas_depth, bd[i]);
    if (td[d->tlas_nodes_len - 1] + blas_depth + 2u > kMaxStack) return ORN_ERROR_BVH_TOO_DEEP;
    std::vector<uint32_t> memo(d->blas_nodes_len, 0xffffffffu);
    const uint32_t slots = std::max(1u, pending2(d->tlas_nodes, static_cast<uint32_t>(d->tlas_nodes_len - 1),
                                                 d->blas_nodes, memo, true));
    if (slots > kMaxStack) return ORN_ERROR_BVH_TOO_DEEP;
    *stack_slots = slots;
    return ORN_OK;
}


orn_tuning default_tuning()
{
    orn_tuning t{};
    t.regroup_lanes = 48;
    t.workgroups_per_cu = 0;  // as many as can be resident at once
    t.bvh_width = 2;
    t.leaf_regroup_lanes = 16;
    t.samples_per_job = 16;
    t.split_percent = 0;
    t.node_steps = 4;
    return t;
}

bool tuning_valid(const orn_tuning *t)
{
    return t && t->regroup_lanes >= 1 && t->regroup_lanes <= 64 && t->workgroups_per_cu <= 64 &&
           (t->bvh_width == 2u || t->bvh_width == 4u) && t->leaf_regroup_lanes >= 1 &&
           t->leaf_regroup_lanes <= 64 && t->samples_per_job <= 0xffffu && t->split_percent <= 100u &&
           t->node_steps <= 16u && t->workgroup_pool <= 1u && t->cost_order <= 1u;
}

// node_steps 0 stands for the default
orn_tuning normalized(orn_tuning t)
{
    if (t.node_steps == 0u) t.node_steps = default_tuning().node_steps;
    return t;
}

// Jobs of one pixel's render: `chunk` samples each, `chunks` of them. A render
// is cut only when current_iteration + k stays exact in f32 for every k up
// to its sample count (a whole number, below 2^24 at the end), and the job
// counter of a band fits 31 bits.
// Whether a render of `iterations` samples from current_iteration can be cut
// (into jobs, or into launches) without changing a bit: the counter is a
// whole number and stays below 2^24, where current_iteration + k is exact.
bool splits_exactly(float current_iteration, uint32_t iterations)
{
    const uint32_t spp = iterations > 1u ? iterations : 1u;
    return current_iteration >= 0.0f && std::floor(current_iteration) == current_iteration &&
           double(current_iteration) + spp <= 16777216.0;
}

void plan_jobs(uint32_t iterations, float current_iteration, uint32_t samples_per_job, uint64_t band_pixels,
               uint32_t *chunk, uint32_t *chunks)
{
    const uint32_t spp = iterations > 1u ? iterations : 1u;
    const bool exact = splits_exactly(current_iteration, iterations);
    uint32_t c = (samples_per_job == 0u || !exact) ? spp : std::min(samples_per_job, spp);
    while (true) {
        const uint32_t n = (spp + c - 1u) / c;
        if (n <= 0xffffu && uint64_t(n) * std::max<uint64_t>(band_pixels, 1) < (1ull << 31)) {
            *chunk = c;
            *chunks = n;
            return;
        }
        c = c >= spp / 2u ? spp : c * 2u;
    }
}

// Buffers and sizes of one render launch.
struct LaunchArgs {
    KernelParams kp;
    const BatchItem *items;
    const BatchTiles *work;
    float4 *acc;
    uint32_t *rng;
    float4 *fb;
    uint32_t *stage;
    uint32_t *queues;
    uint32_t total_tiles;
    bool pool;  // tuning workgroup_pool (honoured for spheres-only scenes, BVH2, one image)
};

// One instantiation per combination of scene features.
// Persistent grid: workgroups_per_cu per compute unit (0: as many as can be
// resident at once), a multiple of the XCD count, at most one per tile.
template <uint32_t F, uint32_t W, bool BATCH, bool POOL>
hipError_t launch_render_k(uint32_t per_cu, uint32_t cus, size_t lds, hipStream_t st, const DevScene &ds,
                           const LaunchArgs &a)
{
    if (per_cu == 0) {
        int blocks = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, render_kernel<F, W, BATCH, POOL>, kBlock, lds) !=
                hipSuccess ||
            blocks < 1) {
            (void)hipGetLastError();  // the launch below reports its own errors
            blocks = 1;
        }
        per_cu = static_cast<uint32_t>(blocks);
    }
    uint32_t grid = cus * per_cu;
    grid = std::min(grid, blocks_for(a.total_tiles, kXcds) * kXcds);
    grid = std::max(kXcds, grid / kXcds * kXcds);
    render_kernel<F, W, BATCH, POOL>
        <<<grid, kBlock, lds, st>>>(ds, a.kp, a.items, a.work, a.acc, a.rng, a.fb, a.stage, a.queues);
    return hipGetLastError();
}

template <uint32_t F, uint32_t W, bool BATCH>
hipError_t launch_render_f(uint32_t per_cu, uint32_t cus, size_t lds, hipStream_t st, const DevScene &ds,
                           const LaunchArgs &a)
{
    if constexpr (!(F & kHasMeshes) && W == 2u && !BATCH)
        if (a.pool) return launch_render_k<F, W, BATCH, true>(per_cu, cus, lds, st, ds, a);
    return launch_render_k<F, W, BATCH, false>(per_cu, cus, lds, st, ds, a);
}

template <bool BATCH>
hipError_t launch_render(uint32_t features, uint32_t width, uint32_t per_cu, uint32_t cus, size_t lds,
                         hipStream_t st, const DevScene &ds, const LaunchArgs &a)
{
#define ORN_CASE(F)                                                                      \
    case F:                                                                              \
        return width == 4u ? launch_render_f<F, 4, BATCH>(per_cu, cus, lds, st, ds, a)  \
                           : launch_render_f<F, 2, BATCH>(per_cu, cus, lds, st, ds, a);
    switch (features & 7u) {
        ORN_CASE(0u)
        ORN_CASE(1u)
        ORN_CASE(2u)
        ORN_CASE(3u)
        ORN_CASE(4u)
        ORN_CASE(5u)
        ORN_CASE(6u)
        ORN_CASE(7u)
    }
#undef ORN_CASE
    return hipErrorInvalidValue;
}

}  // namespace

// Scratch of orn_render_noalloc: the band queue counters, then one stage
// word per pixel. Each device holds its own copy for the life of the
// process. Host code reaches it only through the address
// hipGetSymbolAddress returns for the device (kept in g_scratch); the
// array's name is not a device address on the host.
constexpr uint32_t kScratchQueueWords = 64;
static __device__ __align__(256) uint32_t scratch[kScratchQueueWords + ORN_RENDER_SCRATCH_PIXELS];

struct ScratchSlot {
    uint32_t *ptr;   // this device's scratch (device address)
    uint32_t epoch;  // tag of the latest orn_render_noalloc launch on the device
};
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mutex;
static ScratchSlot g_scratch[kMaxDevices];

// Called with `device` current (orn_device_open), once per device.
static hipError_t init_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch[device].ptr) return hipSuccess;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[device].ptr = static_cast<uint32_t *>(p);
    g_scratch[device].epoch = 0;  // the array starts zeroed: no stage word carries a tag
    return hipSuccess;
}

extern "C" {

const char *orn_status_string(orn_status s)
{
    switch (s) {
    case ORN_OK: return "Success";
    case ORN_ERROR_EMPTY_SCENE: return "EmptyScene";
    case ORN_ERROR_EMPTY_BVH: return "EmptyBvh";
    case ORN_ERROR_INVALID_SCENE: return "InvalidScene";
    case ORN_ERROR_BVH_TOO_DEEP: return "BvhTooDeep";
    case ORN_ERROR_UNSUPPORTED_TEXTURE: return "UnsupportedTexture";
    case ORN_ERROR_INVALID_HANDLE: return "InvalidHandle";
    case ORN_ERROR_SIZE_MISMATCH: return "SizeMismatch";
    case ORN_ERROR_IO: return "IoError";
    case ORN_ERROR_INVALID_FILE: return "InvalidFile";
    case ORN_ERROR_SCRATCH_TOO_SMALL: return "ScratchTooSmall";
    default: return hipGetErrorName(static_cast<hipError_t>(s));
    }
}

orn_status orn_device_count(int *count)
{
    if (!count) return ORN_ERROR_INVALID_VALUE;
    *count = 0;
    hipError_t e = hipGetDeviceCount(count);
    if (e == hipErrorNoDevice) {
        *count = 0;
        return ORN_OK;
    }
    return from_hip(e);
}

orn_status orn_device_open(int device, orn_ctx **out)
{
    if (!out) return ORN_ERROR_INVALID_VALUE;
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) return ORN_ERROR_NO_DEVICE;
    if (device < 0) device = n - 1;
    if (device >= n) return ORN_ERROR_INVALID_DEVICE;
    ORN_TRY(hipSetDevice(device));
    orn_ctx *c = new (std::nothrow) orn_ctx();
    if (!c) return ORN_ERROR_OUT_OF_MEMORY;
    c->device = device;
    c->refs = 1;
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) {
        delete c;
        return from_hip(e);
    }
    if ((e = init_scratch(device)) != hipSuccess) {
        delete c;
        return from_hip(e);
    }
    std::snprintf(c->name, sizeof(c->name), "%s (%s)", prop.name, prop.gcnArchName);
    std::snprintf(c->arch, sizeof(c->arch), "%s", prop.gcnArchName);
    c->compute_units = prop.multiProcessorCount;
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return from_hip(e);
    }
    *out = c;
    return ORN_OK;
}

orn_status orn_device_close(orn_ctx *ctx)
{
    if (!ctx) return ORN_ERROR_INVALID_HANDLE;
    release_ctx(ctx);
    return ORN_OK;
}

orn_status orn_device_info(orn_ctx *ctx, char *name, size_t name_len, int *compute_units)
{
    if (!ctx) return ORN_ERROR_INVALID_HANDLE;
    if (name && name_len) {
        std::strncpy(name, ctx->name, name_len - 1);
        name[name_len - 1] = 0;
    }
    if (compute_units) *compute_units = ctx->compute_units;
    return ORN_OK;
}

}  // extern "C"

namespace {

// Uploads validated arrays. `roots` are the TLAS root indices of the scenes
// the arrays hold (one, or one per image of a batch); their BVH2 and BVH4
// root references are returned in root2/root4. `slots` is the largest BVH2
// stack bound of those scenes.
orn_status upload_scene(orn_ctx *ctx, const orn_scene_desc *desc, uint32_t slots, const uint32_t *roots,
                        size_t n_roots, uint32_t *root2, uint32_t *root4, orn_device_scene **out)
{
    orn_status st = ORN_OK;
    ORN_TRY(hipSetDevice(ctx->device));
    orn_device_scene *s = new (std::nothrow) orn_device_scene();
    if (!s) return ORN_ERROR_OUT_OF_MEMORY;
    s->ctx = ctx;
    ctx->refs++;
    s->refs = 1;
    s->stack_slots = slots;
    s->blas_len = static_cast<uint32_t>(desc->blas_nodes_len);
    s->trace_width = 2u;
    for (int a = 0; a < 3; a++) {  // box of all scenes' TLAS roots
        s->bounds_lo[a] = INFINITY;
        s->bounds_hi[a] = -INFINITY;
    }
    for (size_t r = 0; r < n_roots; r++) {
        const orn_bvh_node &nd = desc->tlas_nodes[roots[r]];
        for (int a = 0; a < 3; a++) {
            const float lo = nd.node_type == 0u ? std::fmin(nd.lmin_or_v0[a], nd.rmin_or_v2[a]) : nd.lmin_or_v0[a];
            const float hi = nd.node_type == 0u ? std::fmax(nd.lmax_or_v1[a], nd.rmax_or_v3[a]) : nd.lmax_or_v1[a];
            s->bounds_lo[a] = std::fmin(s->bounds_lo[a], lo);
            s->bounds_hi[a] = std::fmax(s->bounds_hi[a], hi);
        }
    }
    s->features = desc->textures_len ? kHasTextures : 0u;
    for (size_t i = 0; i < desc->tlas_nodes_len; i++) {
        if (desc->tlas_nodes[i].node_type == 1u) s->features |= kHasSpheres;
        if (desc->tlas_nodes[i].node_type == 2u) s->features |= kHasMeshes;
    }
    DevScene &ds = s->ds;
    auto fail = [&](orn_status e) {
        release_scene(s);
        return e;
    };
    const float4 *f4 = nullptr;
    if ((st = upload(s, reinterpret_cast<const float4 *>(desc->tlas_nodes), 4 * desc->tlas_nodes_len, &f4))) return fail(st);
    ds.tlas = f4;
    if ((st = upload(s, reinterpret_cast<const float4 *>(desc->blas_nodes), 4 * desc->blas_nodes_len, &f4))) return fail(st);
    ds.blas = f4;
    if ((st = upload(s, reinterpret_cast<const float4 *>(desc->normals), desc->normals_len, &f4))) return fail(st);
    ds.normals = f4;
    if ((st = upload(s, reinterpret_cast<const float4 *>(desc->transforms), 4 * desc->transforms_len, &f4))) return fail(st);
    ds.transforms = f4;
    const uint32_t *u32 = nullptr;
    if ((st = upload(s, desc->normal_indices, desc->normal_indices_len, &u32))) return fail(st);
    ds.normal_indices = u32;
    if ((st = upload(s, desc->uv_indices, desc->uv_indices_len, &u32))) return fail(st);
    ds.uv_indices = u32;
    const float2 *f2 = nullptr;
    if ((st = upload(s, reinterpret_cast<const float2 *>(desc->uvs), desc->uvs_len, &f2))) return fail(st);
    ds.uvs = f2;
    const uint4 *u4 = nullptr;
    if ((st = upload(s, reinterpret_cast<const uint4 *>(desc->materials), 2 * desc->materials_len, &u4))) return fail(st);
    ds.materials = u4;
    // BVH2 copies whose child references flag leaves (kLeafRef)
    {
        auto flag = [](const orn_bvh_node *arr, uint32_t i) { return arr[i].node_type != 0u ? (i | kLeafRef) : i; };
        std::vector<orn_bvh_node> t2(desc->tlas_nodes, desc->tlas_nodes + desc->tlas_nodes_len);
        std::vector<orn_bvh_node> b2(desc->blas_nodes, desc->blas_nodes + desc->blas_nodes_len);
        for (orn_bvh_node &nd : t2) {
            if (nd.node_type == 0u) {
                nd.left_or_custom_id = flag(desc->tlas_nodes, nd.left_or_custom_id);
                nd.right_or_material_index = flag(desc->tlas_nodes, nd.right_or_material_index);
            } else if (nd.node_type == 2u) {
                nd.left_or_custom_id = flag(desc->blas_nodes, nd.left_or_custom_id);
            }
        }
        for (orn_bvh_node &nd : b2)
            if (nd.node_type == 0u) {
                nd.left_or_custom_id = flag(desc->blas_nodes, nd.left_or_custom_id);
                nd.right_or_material_index = flag(desc->blas_nodes, nd.right_or_material_index);
            }
        for (size_t i = 0; i < n_roots; i++) root2[i] = flag(desc->tlas_nodes, roots[i]);
        ds.tlas2_root = root2[0];
        if ((st = upload(s, reinterpret_cast<const float4 *>(t2.data()), 4 * t2.size(), &f4))) return fail(st);
        ds.tlas2 = f4;
        if ((st = upload(s, reinterpret_cast<const float4 *>(b2.data()), 4 * b2.size(), &f4))) return fail(st);
        ds.blas2 = f4;
    }
    // derived BVH4 arrays and the TLAS copy whose mesh leaves point at
    // their BLAS's BVH4 root
    {
        Bvh4Builder t4(desc->tlas_nodes, desc->tlas_nodes_len);
        for (size_t i = 0; i < n_roots; i++) root4[i] = t4.build(roots[i]);
        Bvh4Builder b4(desc->blas_nodes, desc->blas_nodes_len);
        std::vector<uint32_t> blas_ref(desc->blas_nodes_len, 0u);
        std::vector<orn_bvh_node> leaves(desc->tlas_nodes, desc->tlas_nodes + desc->tlas_nodes_len);
        for (orn_bvh_node &nd : leaves)
            if (nd.node_type == 2u) {
                const uint32_t r = nd.left_or_custom_id;
                blas_ref[r] = b4.build(r);
                nd.left_or_custom_id = blas_ref[r];
            }
        std::vector<uint32_t> memo(desc->blas_nodes_len, 0xffffffffu);
        uint32_t slots4 = 1;
        for (size_t i = 0; i < n_roots; i++)
            slots4 = std::max(slots4, pending4(t4.out, root4[i], desc->tlas_nodes, &b4.out, &memo, blas_ref.data()));
        if (slots4 > kMaxStack) return fail(ORN_ERROR_BVH_TOO_DEEP);
        s->stack_slots4 = slots4;
        ds.tlas4_root = root4[0];
        if ((st = upload(s, reinterpret_cast<const float4 *>(t4.out.data()), 8 * t4.out.size(), &f4))) return fail(st);
        ds.tlas4 = f4;
        if ((st = upload(s, reinterpret_cast<const float4 *>(b4.out.data()), 8 * b4.out.size(), &f4))) return fail(st);
        ds.blas4 = f4;
        if ((st = upload(s, reinterpret_cast<const float4 *>(leaves.data()), 4 * leaves.size(), &f4))) return fail(st);
        ds.tlas_leaves = f4;
    }
    // textures: one texel pool with 16-byte aligned offsets
    std::vector<DevTexture> tex(desc->textures_len);
    uint64_t pool = 0;
    for (size_t i = 0; i < desc->textures_len; i++) {
        const orn_texture_desc &t = desc->textures[i];
        tex[i].offset = pool;
        tex[i].width = t.width;
        tex[i].height = t.height;
        tex[i].components = t.num_components;
        tex[i].hdr = t.is_hdr ? 1u : 0u;
        const uint64_t bytes = uint64_t(t.width) * t.height * t.num_components * t.bytes_per_component;
        pool += (bytes + 15u) & ~uint64_t(15);
    }
    std::vector<uint8_t> texels(pool);
    for (size_t i = 0; i < desc->textures_len; i++) {
        const orn_texture_desc &t = desc->textures[i];
        const uint64_t bytes = uint64_t(t.width) * t.height * t.num_components * t.bytes_per_component;
        std::memcpy(texels.data() + tex[i].offset, t.data, bytes);
    }
    const DevTexture *dt = nullptr;
    if ((st = upload(s, tex.data(), tex.size(), &dt))) return fail(st);
    ds.textures = dt;
    const uint8_t *u8 = nullptr;
    if ((st = upload(s, texels.data(), texels.size(), &u8))) return fail(st);
    ds.texels = u8;
    ds.tlas_len = static_cast<uint32_t>(desc->tlas_nodes_len);
    ds.textures_len = static_cast<uint32_t>(desc->textures_len);
    ds.stack_slots = slots;
    *out = s;
    return ORN_OK;
}

}  // namespace

extern "C" {

orn_status orn_scene_upload(orn_ctx *ctx, const orn_scene_desc *desc, orn_device_scene **out)
{
    if (!ctx) return ORN_ERROR_INVALID_HANDLE;
    if (!desc || !out) return ORN_ERROR_INVALID_VALUE;
    *out = nullptr;
    uint32_t slots = 0;
    orn_status st = validate(desc, &slots);
    if (st != ORN_OK) return st;
    const uint32_t root = static_cast<uint32_t>(desc->tlas_nodes_len - 1);
    uint32_t root2 = 0, root4 = 0;
    return upload_scene(ctx, desc, slots, &root, 1, &root2, &root4, out);
}

orn_status orn_scene_validate(const orn_scene_desc *desc, uint32_t *stack_slots)
{
    if (!desc) return ORN_ERROR_INVALID_VALUE;
    uint32_t slots = 0;
    orn_status st = validate(desc, &slots);
    if (st == ORN_OK && stack_slots) *stack_slots = slots;
    return st;
}

orn_status orn_scene_destroy(orn_device_scene *s)
{
    if (!s) return ORN_ERROR_INVALID_HANDLE;
    release_scene(s);
    return ORN_OK;
}

orn_status orn_scene_stack_depth(orn_device_scene *s, uint32_t *depth)
{
    if (!s) return ORN_ERROR_INVALID_HANDLE;
    if (!depth) return ORN_ERROR_INVALID_VALUE;
    *depth = s->trace_width == 4u ? s->stack_slots4 : s->stack_slots;
    return ORN_OK;
}

orn_status orn_scene_set_trace_width(orn_device_scene *s, uint32_t width)
{
    if (!s) return ORN_ERROR_INVALID_HANDLE;
    if (width != 2u && width != 4u) return ORN_ERROR_INVALID_VALUE;
    s->trace_width = width;
    return ORN_OK;
}

orn_status orn_target_create(orn_device_scene *s, uint32_t width, uint32_t height, uint32_t seed,
                             orn_target **out)
{
    if (!s) return ORN_ERROR_INVALID_HANDLE;
    if (!out || width == 0 || height == 0) return ORN_ERROR_INVALID_VALUE;
    const uint64_t n = uint64_t(width) * height;
    if (n > kMaxPixels) return ORN_ERROR_INVALID_VALUE;
    *out = nullptr;
    ORN_TRY(hipSetDevice(s->ctx->device));
    orn_target *t = new (std::nothrow) orn_target();
    if (!t) return ORN_ERROR_OUT_OF_MEMORY;
    t->scene = s;
    s->refs++;
    t->tuning = default_tuning();
    t->width = width;
    t->height = height;
    auto fail = [&](hipError_t e) {
        orn_target_destroy(t);
        return from_hip(e);
    };
    hipError_t e;
    if ((e = hipMalloc(&t->fb, n * sizeof(float4))) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&t->acc, n * sizeof(float4))) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&t->rng, n * sizeof(uint32_t))) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&t->queues, 64 * sizeof(uint32_t))) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&t->stage, n * sizeof(uint32_t))) != hipSuccess) return fail(e);
    if ((e = hipEventCreate(&t->ev_start)) != hipSuccess) return fail(e);
    if ((e = hipEventCreate(&t->ev_stop)) != hipSuccess) return fail(e);
    hipStream_t st = s->ctx->stream;
    const uint32_t nn = static_cast<uint32_t>(n);
    zero_kernel<<<blocks_for(nn, 256), 256, 0, st>>>(t->fb, t->acc, nn);
    seed_rng_kernel<<<blocks_for(nn, 256), 256, 0, st>>>(t->rng, nn, seed * nn);
    fill_u32_kernel<<<blocks_for(nn, 256), 256, 0, st>>>(t->stage, nn, 0u);
    if ((e = hipGetLastError()) != hipSuccess) return fail(e);
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return fail(e);
    *out = t;
    return ORN_OK;
}

orn_status orn_target_destroy(orn_target *t)
{
    if (!t) return ORN_ERROR_INVALID_HANDLE;
    (void)set_device(t->scene->ctx);
    (void)hipStreamSynchronize(t->scene->ctx->stream);
    if (t->fb) (void)hipFree(t->fb);
    if (t->acc) (void)hipFree(t->acc);
    if (t->rng) (void)hipFree(t->rng);
    if (t->queues) (void)hipFree(t->queues);
    if (t->stage) (void)hipFree(t->stage);
    if (t->cost) (void)hipFree(t->cost);
    if (t->order) (void)hipFree(t->order);
    if (t->counts) (void)hipFree(t->counts);
    if (t->totals) (void)hipFree(t->totals);
    if (t->ev_start) (void)hipEventDestroy(t->ev_start);
    if (t->ev_stop) (void)hipEventDestroy(t->ev_stop);
    orn_device_scene *s = t->scene;
    delete t;
    release_scene(s);
    return ORN_OK;
}

}  // extern "C"

namespace {

// The render launch of one image on buffers the caller provides: the
// queues are reset and the kernel enqueued on `st`. `epoch` tags this
// launch's stage words (unique among the launches that used `stage` since it
// was last cleared). Allocates nothing.
hipError_t enqueue_render(orn_device_scene *s, const orn_constant_params *params, uint32_t iterations,
                          const orn_tuning &tuning, float4 *acc, uint32_t *rng, float4 *fb, uint32_t *stage,
                          uint32_t *queues, uint32_t epoch, hipStream_t st, const uint32_t *order = nullptr,
                          uint32_t *cost = nullptr)
{
    KernelParams kp{};
    kp.p = *params;
    kp.iterations = iterations;
    kp.tiles_x = blocks_for(params->width, kTile);
    kp.tiles = kp.tiles_x * blocks_for(params->height, kTile);
    kp.tiles_per_xcd = blocks_for(kp.tiles, kXcds);
    kp.regroup = tuning.regroup_lanes;
    kp.leaf_regroup = tuning.leaf_regroup_lanes;
    kp.node_steps = tuning.node_steps;
    const uint32_t width = tuning.bvh_width == 2u ? 2u : 4u;
    kp.root = width == 4u ? s->ds.tlas4_root : s->ds.tlas2_root;
    plan_jobs(iterations, params->current_iteration, tuning.samples_per_job,
              uint64_t(kp.tiles_per_xcd) * kTile * kTile, &kp.chunk, &kp.chunks);
    DevScene ds = s->ds;
    ds.stack_slots = width == 4u ? s->stack_slots4 : s->stack_slots;
    const size_t lds = size_t(ds.stack_slots) * kBlock * sizeof(uint32_t);
    kp.epoch = epoch;
    kp.split_pct = tuning.split_percent;
    kp.pixels = params->width * params->height;
    kp.order = order;
    kp.cost = cost;
    LaunchArgs a{kp, nullptr, nullptr, acc, rng, fb, stage, queues, kp.tiles, tuning.workgroup_pool != 0u};
    reset_queues_kernel<<<1, 64, 0, st>>>(queues);
    return launch_render<false>(s->features, width, tuning.workgroups_per_cu,
                                static_cast<uint32_t>(s->ctx->compute_units), lds, st, ds, a);
}

// Buffers of a target's cost-ordered queues, allocated on first use; the
// costs start at zero.
hipError_t ensure_cost_order(orn_target *t, hipStream_t st)
{
    if (t->cost) return hipSuccess;
    const uint32_t n = t->width * t->height;
    const uint32_t tiles = blocks_for(t->width, kTile) * blocks_for(t->height, kTile);
    const uint32_t positions = blocks_for(tiles, kXcds) * kXcds * (kTile * kTile);
    hipError_t e = hipMalloc(&t->cost, size_t(n) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&t->order, size_t(positions) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&t->counts, size_t(tiles) * kCostClasses * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&t->totals, size_t(kXcds) * kCostClasses * sizeof(uint32_t));
    if (e != hipSuccess) {
        (void)hipFree(t->cost);
        (void)hipFree(t->order);
        (void)hipFree(t->counts);
        (void)hipFree(t->totals);
        t->cost = t->order = t->counts = t->totals = nullptr;
        return e;
    }
    fill_u32_kernel<<<blocks_for(n, 256), 256, 0, st>>>(t->cost, n, 0u);
    t->costs_recorded = false;
    return hipGetLastError();
}

// The queue order of a target's next render from the costs its previous
// recording render left (and the costs cleared for the next recording).
void enqueue_queue_order(orn_target *t, hipStream_t st)
{
    OrderShape o{};
    o.tiles_x = blocks_for(t->width, kTile);
    o.tiles = o.tiles_x * blocks_for(t->height, kTile);
    o.tiles_per_xcd = blocks_for(o.tiles, kXcds);
    o.width = t->width;
    o.height = t->height;
    enqueue_order_kernels(o, o.tiles, t->cost, t->counts, t->totals, t->order, st);
}

}  // namespace

extern "C" {

orn_status orn_render_async(orn_target *t, const orn_constant_params *params, uint32_t iterations)
{
    if (!t) return ORN_ERROR_INVALID_HANDLE;
    if (!params) return ORN_ERROR_INVALID_VALUE;
    if (params->width != t->width || params->height != t->height) return ORN_ERROR_SIZE_MISMATCH;
    orn_device_scene *s = t->scene;
    ORN_TRY(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    const uint32_t n = t->width * t->height;
    auto next_epoch = [&]() {
        if (++t->epoch > 0xffffu) {  // tags wrap: clear the stage words
            fill_u32_kernel<<<blocks_for(n, 256), 256, 0, st>>>(t->stage, n, 0u);
            t->epoch = 1;
        }
    };
    next_epoch();
    const bool by_cost = t->tuning.cost_order != 0u;
    if (by_cost) ORN_TRY(ensure_cost_order(t, st));
    ORN_TRY(hipEventRecord(t->ev_start, st));
    const uint32_t *order = nullptr;
    if (by_cost && t->costs_recorded) {
        enqueue_queue_order(t, st);
        order = t->order;
    } else if (by_cost && splits_exactly(params->current_iteration, iterations) && iterations >= 2u) {
        // No costs recorded yet: the first eighth of the samples runs in
        // tile order and records them, and the rest, queued by those costs,
        // continues it (a render of k then n - k samples from
        // current_iteration + k gives the same bits as one of n).
        const uint32_t first = std::max(1u, iterations / 8u);
        ORN_TRY(enqueue_render(s, params, first, t->tuning, t->acc, t->rng, t->fb, t->stage, t->queues, t->epoch,
                               st, nullptr, t->cost));
        next_epoch();
        enqueue_queue_order(t, st);
        orn_constant_params rest = *params;
        rest.current_iteration = params->current_iteration + static_cast<float>(first);
        ORN_TRY(enqueue_render(s, &rest, iterations - first, t->tuning, t->acc, t->rng, t->fb, t->stage, t->queues,
                               t->epoch, st, t->order, t->cost));
        ORN_TRY(hipEventRecord(t->ev_stop, st));
        t->timed = true;
        t->costs_recorded = true;
        return ORN_OK;
    }
    ORN_TRY(enqueue_render(s, params, iterations, t->tuning, t->acc, t->rng, t->fb, t->stage, t->queues, t->epoch,
                           st, order, by_cost ? t->cost : nullptr));
    ORN_TRY(hipEventRecord(t->ev_stop, st));
    t->timed = true;
    if (by_cost) t->costs_recorded = true;
    return ORN_OK;
}

orn_status orn_render_noalloc(orn_device_scene *s, const orn_constant_params *params, uint32_t iterations,
                              const orn_tuning *tuning, float *accumulation, uint32_t *rng, float *framebuffer,
                              void *stream)
{
    if (!s) return ORN_ERROR_INVALID_HANDLE;
    if (!params || !accumulation || !rng || !framebuffer) return ORN_ERROR_INVALID_VALUE;
    if (params->width == 0 || params->height == 0) return ORN_ERROR_INVALID_VALUE;
    if (uint64_t(params->width) * params->height > ORN_RENDER_SCRATCH_PIXELS) return ORN_ERROR_SCRATCH_TOO_SMALL;
    if (tuning && !tuning_valid(tuning)) return ORN_ERROR_INVALID_VALUE;
    const orn_tuning tn = tuning ? normalized(*tuning) : default_tuning();
    const int device = s->ctx->device;
    if (device < 0 || device >= kMaxDevices) return ORN_ERROR_INVALID_DEVICE;
    ORN_TRY(hipSetDevice(device));
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : s->ctx->stream;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    ScratchSlot &sl = g_scratch[device];
    if (!sl.ptr) return ORN_ERROR_INVALID_DEVICE;  // set when the device was opened
    uint32_t *queues = sl.ptr;
    uint32_t *stage = sl.ptr + kScratchQueueWords;
    if (++sl.epoch > 0xffffu) {  // tags wrap: clear every stage word an earlier call may have left
        fill_u32_kernel<<<blocks_for(ORN_RENDER_SCRATCH_PIXELS, 256), 256, 0, st>>>(stage, ORN_RENDER_SCRATCH_PIXELS,
                                                                                   0u);
        sl.epoch = 1;
    }
    ORN_TRY(enqueue_render(s, params, iterations, tn, reinterpret_cast<float4 *>(accumulation), rng,
                           reinterpret_cast<float4 *>(framebuffer), stage, queues, sl.epoch, st));
    return ORN_OK;
}

orn_status orn_target_wait(orn_target *t)
{
    if (!t) return ORN_ERROR_INVALID_HANDLE;
    ORN_TRY(hipSetDevice(t->scene->ctx->device));
    ORN_TRY(hipStreamSynchronize(t->scene->ctx->stream));
    return ORN_OK;
}

orn_status orn_render(orn_target *t, const orn_constant_params *params, uint32_t iterations)
{
    orn_status s = orn_render_async(t, params, iterations);
    if (s != ORN_OK) return s;
    return orn_target_wait(t);
}

orn_status orn_target_set_tuning(orn_target *t, const orn_tuning *tuning)
{
    if (!t) return ORN_ERROR_INVALID_HANDLE;
    if (!tuning_valid(tuning)) return ORN_ERROR_INVALID_VALUE;
    t->tuning = normalized(*tuning);
    return ORN_OK;
}

orn_status orn_target_get_tuning(orn_target *t, orn_tuning *tuning)
{
    if (!t) return ORN_ERROR_INVALID_HANDLE;
    if (!tuning) return ORN_ERROR_INVALID_VALUE;
    *tuning = t->tuning;
    return ORN_OK;
}

orn_status orn_target_last_render_ms(orn_target *t, float *ms)
{
    if (!t) return ORN_ERROR_INVALID_HANDLE;
    if (!ms) return ORN_ERROR_INVALID_VALUE;
    *ms = 0.0f;
    if (!t->timed) return ORN_OK;
    ORN_TRY(hipEventSynchronize(t->ev_stop));
    ORN_TRY(hipEventElapsedTime(ms, t->ev_start, t->ev_stop));
    return ORN_OK;
}

static orn_status readback(orn_target *t, void *dst, const void *src, size_t elem, size_t pixels)
{
    if (!t) return ORN_ERROR_INVALID_HANDLE;
    if (!dst) return ORN_ERROR_INVALID_VALUE;
    const size_t n = size_t(t->width) * t->height;
    if (pixels < n) return ORN_ERROR_SIZE_MISMATCH;
    ORN_TRY(hipSetDevice(t->scene->ctx->device));
    hipStream_t st = t->scene->ctx->stream;
    ORN_TRY(hipMemcpyAsync(dst, src, n * elem, hipMemcpyDeviceToHost, st));
    ORN_TRY(hipStreamSynchronize(st));
    return ORN_OK;
}

orn_status orn_get_framebuffer(orn_target *t, float *dst, size_t pixels)
{
    return readback(t, dst, t ? t->fb : nullptr, sizeof(float4), pixels);
}

orn_status orn_get_accumulation(orn_target *t, float *dst, size_t pixels)
{
    return readback(t, dst, t ? t->acc : nullptr, sizeof(float4), pixels);
}

orn_status orn_get_rng_state(orn_target *t, uint32_t *dst, size_t pixels)
{
    return readback(t, dst, t ? t->rng : nullptr, sizeof(uint32_t), pixels);
}

orn_status orn_get_pixel_costs(orn_target *t, uint32_t *dst, size_t pixels)
{
    if (t && !t->costs_recorded) return ORN_ERROR_INVALID_VALUE;
    return readback(t, dst, t ? t->cost : nullptr, sizeof(uint32_t), pixels);
}

orn_status orn_set_state(orn_target *t, const float *accumulation, const uint32_t *rng, size_t pixels)
{
    if (!t) return ORN_ERROR_INVALID_HANDLE;
    const size_t n = size_t(t->width) * t->height;
    if (pixels != n) return ORN_ERROR_SIZE_MISMATCH;
    ORN_TRY(hipSetDevice(t->scene->ctx->device));
    hipStream_t st = t->scene->ctx->stream;
    if (accumulation) ORN_TRY(hipMemcpyAsync(t->acc, accumulation, n * sizeof(float4), hipMemcpyHostToDevice, st));
    if (rng) ORN_TRY(hipMemcpyAsync(t->rng, rng, n * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    ORN_TRY(hipStreamSynchronize(st));
    return ORN_OK;
}

static orn_status trace_rays(orn_device_scene *s, float t_min, const float *rays, size_t n, orn_hit *out,
                             bool packet)
{
    if (!s) return ORN_ERROR_INVALID_HANDLE;
    if ((!rays || !out) && n) return ORN_ERROR_INVALID_VALUE;
    if (n == 0) return ORN_OK;
    if (n > 0xffffffffull) return ORN_ERROR_INVALID_VALUE;
    ORN_TRY(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    float *d_rays = nullptr;
    orn_hit *d_out = nullptr;
    hipError_t e = hipMalloc(&d_rays, n * 6 * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&d_out, n * sizeof(orn_hit));
    if (e == hipSuccess) e = hipMemcpyAsync(d_rays, rays, n * 6 * sizeof(float), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        DevScene ds = s->ds;
        ds.stack_slots = s->trace_width == 4u ? s->stack_slots4 : s->stack_slots;
        const size_t lds = size_t(ds.stack_slots) * kBlock * sizeof(uint32_t);
        const uint32_t blocks = blocks_for(static_cast<uint32_t>(n), kBlock);
        if (packet) {
            ds.stack_slots = s->stack_slots;
            trace_packet_kernel<<<blocks, kBlock, 0, st>>>(ds, t_min, d_rays, static_cast<uint32_t>(n), d_out);
        } else if (s->trace_width == 4u) {
            trace_kernel<4><<<blocks, kBlock, lds, st>>>(ds, t_min, d_rays, static_cast<uint32_t>(n), d_out);
        } else {
            trace_kernel<2><<<blocks, kBlock, lds, st>>>(ds, t_min, d_rays, static_cast<uint32_t>(n), d_out);
        }
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, n * sizeof(orn_hit), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipStreamSynchronize(st);
    if (d_rays) (void)hipFree(d_rays);
    if (d_out) (void)hipFree(d_out);
    return from_hip(e);
}

orn_status orn_trace_rays(orn_device_scene *s, float t_min, const float *rays, size_t n, orn_hit *out)
{
    return trace_rays(s, t_min, rays, n, out, false);
}

orn_status orn_trace_rays_packet(orn_device_scene *s, float t_min, const float *rays, size_t n, orn_hit *out)
{
    return trace_rays(s, t_min, rays, n, out, true);
}

orn_status orn_eval_math(orn_ctx *ctx, int op, const float *a, const float *b, float *out, size_t n)
{
    if (!ctx) return ORN_ERROR_INVALID_HANDLE;
    if (op < 0 || op > 5 || !a || !out || ((op == 2 || op == 4) && !b)) return ORN_ERROR_INVALID_VALUE;
    if (n == 0) return ORN_OK;
    if (n > 0xffffffffull) return ORN_ERROR_INVALID_VALUE;
    ORN_TRY(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    float *d = nullptr;
    hipError_t e = hipMalloc(&d, 3 * n * sizeof(float));
    if (e == hipSuccess) e = hipMemcpyAsync(d, a, n * sizeof(float), hipMemcpyHostToDevice, st);
    if (e == hipSuccess && b) e = hipMemcpyAsync(d + n, b, n * sizeof(float), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        math_kernel<<<blocks_for(static_cast<uint32_t>(n), 256), 256, 0, st>>>(op, d, d + n, d + 2 * n,
                                                                              static_cast<uint32_t>(n));
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, d + 2 * n, n * sizeof(float), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipStreamSynchronize(st);
    if (d) (void)hipFree(d);
    return from_hip(e);
}

orn_status orn_calibration_copy(orn_ctx *ctx, size_t bytes, float *ms)
{
    if (!ctx) return ORN_ERROR_INVALID_HANDLE;
    if (!ms || bytes == 0 || bytes % 16) return ORN_ERROR_INVALID_VALUE;
    ORN_TRY(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    float4 *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&a, bytes);
    if (e == hipSuccess) e = hipMalloc(&b, bytes);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    const uint64_t n = bytes / 16;
    if (e == hipSuccess) {
        zero_kernel<<<blocks_for(static_cast<uint32_t>(std::min<uint64_t>(n, 1u << 30)), 256), 256, 0, st>>>(
            a, b, static_cast<uint32_t>(std::min<uint64_t>(n, 1u << 30)));
        e = hipEventRecord(e0, st);
    }
    if (e == hipSuccess) {
        copy_kernel<<<ctx->compute_units * 8, 256, 0, st>>>(a, b, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
    (void)hipStreamSynchronize(st);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return from_hip(e);
}

}  // extern "C"

// ===========================================================================
// Batches
// ===========================================================================

namespace {

// The arrays of many scenes concatenated into one, with every index
// rebased: node links, BLAS roots, leaf transform and material ids,
// triangle ids, normal and uv indices, texture ids.
struct MergedScene {
    std::vector<orn_bvh_node> tlas, blas;
    std::vector<float> normals, uvs, transforms;
    std::vector<uint32_t> normal_indices, uv_indices;
    std::vector<orn_material> materials;
    std::vector<orn_texture_desc> textures;
    std::vector<uint32_t> roots;     // TLAS root of each scene
    std::vector<uint32_t> tex_base;  // first texture of each scene
    orn_scene_desc desc{};
};

orn_status merge_scenes(const orn_batch_item *items, size_t n, MergedScene &m, uint32_t *slots)
{
    *slots = 1;
    for (size_t i = 0; i < n; i++) {
        const orn_scene_desc *d = items[i].scene;
        if (!d) return ORN_ERROR_INVALID_VALUE;
        uint32_t sl = 0;
        const orn_status st = validate(d, &sl);
        if (st != ORN_OK) return st;
        *slots = std::max(*slots, sl);
        const uint32_t tlas_base = static_cast<uint32_t>(m.tlas.size());
        const uint32_t blas_base = static_cast<uint32_t>(m.blas.size());
        const uint32_t mat_base = static_cast<uint32_t>(m.materials.size());
        const uint32_t tid_base = static_cast<uint32_t>(m.transforms.size() / 32);  // [inverse, model] pairs
        const uint32_t tri_base = static_cast<uint32_t>(m.normal_indices.size() / 3);
        const uint32_t nrm_base = static_cast<uint32_t>(m.normals.size() / 4);
        const uint32_t uv_base = static_cast<uint32_t>(m.uvs.size() / 2);
        const uint32_t tex_base = static_cast<uint32_t>(m.textures.size());
        if (uint64_t(m.tlas.size()) + d->tlas_nodes_len >= 0x7ffffff0u ||
            uint64_t(m.blas.size()) + d->blas_nodes_len >= 0x7ffffff0u)
            return ORN_ERROR_INVALID_SCENE;
        for (size_t k = 0; k < d->tlas_nodes_len; k++) {
            orn_bvh_node nd = d->tlas_nodes[k];
            if (nd.node_type == 0u) {
                nd.left_or_custom_id += tlas_base;
                nd.right_or_material_index += tlas_base;
            } else {
                if (nd.node_type == 2u) nd.left_or_custom_id += blas_base;
                nd.right_or_material_index += mat_base;
                nd.transform_id += tid_base;
            }
            m.tlas.push_back(nd);
        }
        for (size_t k = 0; k < d->blas_nodes_len; k++) {
            orn_bvh_node nd = d->blas_nodes[k];
            if (nd.node_type == 0u) {
                nd.left_or_custom_id += blas_base;
                nd.right_or_material_index += blas_base;
            } else {
                nd.left_or_custom_id += tri_base;
            }
            m.blas.push_back(nd);
        }
        m.normals.insert(m.normals.end(), d->normals, d->normals + 4 * d->normals_len);
        m.uvs.insert(m.uvs.end(), d->uvs, d->uvs + 2 * d->uvs_len);
        // normal and uv indices share the triangle id (3 per triangle): both
        // are padded to one length, a multiple of 3; padding is never read
        const size_t len = (std::max(d->normal_indices_len, d->uv_indices_len) + 2) / 3 * 3;
        for (size_t k = 0; k < len; k++) {
            m.normal_indices.push_back(k < d->normal_indices_len ? d->normal_indices[k] + nrm_base : 0u);
            m.uv_indices.push_back(k < d->uv_indices_len ? d->uv_indices[k] + uv_base : 0u);
        }
        m.transforms.insert(m.transforms.end(), d->transforms, d->transforms + 16 * d->transforms_len);
        if (d->transforms_len % 2) m.transforms.insert(m.transforms.end(), 16, 0.0f);
        for (size_t k = 0; k < d->materials_len; k++) {
            orn_material mt = d->materials[k];
            mt.albedo_texture_index = mt.albedo_texture_index < d->textures_len
                                          ? mt.albedo_texture_index + tex_base
                                          : 0xffffffffu;
            m.materials.push_back(mt);
        }
        m.textures.insert(m.textures.end(), d->textures, d->textures + d->textures_len);
        m.roots.push_back(static_cast<uint32_t>(tlas_base + d->tlas_nodes_len - 1));
        m.tex_base.push_back(tex_base);
    }
    orn_scene_desc &o = m.desc;
    o.tlas_nodes = m.tlas.data();
    o.tlas_nodes_len = m.tlas.size();
    o.blas_nodes = m.blas.data();
    o.blas_nodes_len = m.blas.size();
    o.normals = m.normals.data();
    o.normals_len = m.normals.size() / 4;
    o.normal_indices = m.normal_indices.data();
    o.normal_indices_len = m.normal_indices.size();
    o.uvs = m.uvs.data();
    o.uvs_len = m.uvs.size() / 2;
    o.uv_indices = m.uv_indices.data();
    o.uv_indices_len = m.uv_indices.size();
    o.transforms = m.transforms.data();
    o.transforms_len = m.transforms.size() / 16;
    o.materials = m.materials.data();
    o.materials_len = m.materials.size();
    o.textures = m.textures.data();
    o.textures_len = m.textures.size();
    return ORN_OK;
}

}  // namespace

struct orn_batch {
    orn_device_scene *scene;
    size_t n;
    std::vector<uint32_t> width, height, tiles_x, tiles, px_base, root2, root4, tex_base;
    uint32_t total_px;
    uint32_t total_tiles;
    BatchItem *h_items;  // pinned staging of the item table, read by the graph's copy
    BatchItem *d_items;
    BatchTiles *h_work;  // pinned staging of the tile work list
    BatchTiles *d_work;
    size_t work_bytes;
    std::vector<uint64_t> work_cost;  // per-image costs the current work list was ordered by
    float4 *fb, *acc;
    uint32_t *rng;
    uint32_t *stage;  // per pixel: launch tag << 16 | jobs finished in that launch
    uint32_t epoch;   // tag of the latest launch
    uint32_t *queues;
    orn_tuning tuning;
    hipGraph_t graph;
    hipGraphExec_t exec;
    bool graph_ok;
    hipEvent_t ev_start, ev_stop;
    bool timed;
    // cost-ordered queues (tuning cost_order), allocated by the first capture
    // that asks for them
    uint32_t *cost = nullptr;    // per pooled pixel: GPU time of its jobs in the latest render
    uint32_t *order = nullptr;   // per work-list position: the band's tile-pixel queued there
    uint32_t *counts = nullptr;  // per work-list tile and cost class
    uint32_t *totals = nullptr;  // per band and cost class
};

namespace {

void free_batch(orn_batch *b)
{
    orn_device_scene *s = b->scene;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->h_items) (void)hipHostFree(b->h_items);
    if (b->h_work) (void)hipHostFree(b->h_work);
    if (b->d_items) (void)hipFree(b->d_items);
    if (b->d_work) (void)hipFree(b->d_work);
    if (b->cost) (void)hipFree(b->cost);
    if (b->order) (void)hipFree(b->order);
    if (b->counts) (void)hipFree(b->counts);
    if (b->totals) (void)hipFree(b->totals);
    if (b->fb) (void)hipFree(b->fb);
    if (b->acc) (void)hipFree(b->acc);
    if (b->rng) (void)hipFree(b->rng);
    if (b->queues) (void)hipFree(b->queues);
    if (b->stage) (void)hipFree(b->stage);
    if (b->ev_start) (void)hipEventDestroy(b->ev_start);
    if (b->ev_stop) (void)hipEventDestroy(b->ev_stop);
    delete b;
    release_scene(s);
}

// Orders the tile work list by the images' expected cost per pixel
// (samples x (depth + 1)), most expensive first, and deals it out to the
// eight bands in runs of 4 tiles, so every band starts on the long pixels
// and all bands hold about the same work. The last entry of `cost` is the
// largest jobs-per-pixel count of the images.
void build_work(orn_batch *b, const std::vector<uint64_t> &cost)
{
    std::vector<uint32_t> order(b->n);  // images, costliest first
    for (size_t i = 0; i < b->n; i++) order[i] = static_cast<uint32_t>(i);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t c) { return cost[a] > cost[c]; });
    std::vector<std::vector<uint2>> band(kXcds);
    constexpr uint32_t kRun = 4;
    uint32_t dealt = 0;
    for (uint32_t i : order)
        for (uint32_t t = 0; t < b->tiles[i]; t++, dealt++) band[(dealt / kRun) % kXcds].push_back(make_uint2(i, t));
    b->h_work->max_chunks = static_cast<uint32_t>(cost[b->n]);
    uint32_t at = 0;
    for (uint32_t k = 0; k < kXcds; k++) {
        b->h_work->band_begin[k] = at;
        b->h_work->band_tiles[k] = static_cast<uint32_t>(band[k].size());
        std::memcpy(&b->h_work->tiles[at], band[k].data(), band[k].size() * sizeof(uint2));
        at += static_cast<uint32_t>(band[k].size());
    }
    b->work_cost = cost;
}

// Buffers of a batch's cost-ordered queues, allocated on first use; the
// costs start at zero (so the first render queues in work-list order).
hipError_t ensure_batch_cost_order(orn_batch *b, hipStream_t st)
{
    if (b->cost) return hipSuccess;
    const size_t positions = size_t(b->total_tiles) * (kTile * kTile);
    hipError_t e = hipMalloc(&b->cost, size_t(b->total_px) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&b->order, positions * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&b->counts, size_t(b->total_tiles) * kCostClasses * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&b->totals, size_t(kXcds) * kCostClasses * sizeof(uint32_t));
    if (e != hipSuccess) {
        (void)hipFree(b->cost);
        (void)hipFree(b->order);
        (void)hipFree(b->counts);
        (void)hipFree(b->totals);
        b->cost = b->order = b->counts = b->totals = nullptr;
        return e;
    }
    fill_u32_kernel<<<blocks_for(b->total_px, 256), 256, 0, st>>>(b->cost, b->total_px, 0u);
    return hipGetLastError();
}

// Captures [item table copy, queue reset, render launch] as one graph; with
// tuning cost_order, the queue-order kernels run between the copy and the
// launch, on the costs the previous replay recorded.
orn_status capture_batch(orn_batch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_ok = false;
    orn_device_scene *s = b->scene;
    hipStream_t st = s->ctx->stream;
    const uint32_t width = b->tuning.bvh_width == 4u ? 4u : 2u;
    DevScene ds = s->ds;
    ds.stack_slots = width == 4u ? s->stack_slots4 : s->stack_slots;
    const size_t lds = size_t(ds.stack_slots) * kBlock * sizeof(uint32_t);
    KernelParams kp{};
    kp.regroup = b->tuning.regroup_lanes;
    kp.leaf_regroup = b->tuning.leaf_regroup_lanes;
    kp.node_steps = b->tuning.node_steps;
    kp.split_pct = b->tuning.split_percent;
    kp.pixels = b->total_px;
    const bool by_cost = b->tuning.cost_order != 0u;
    if (by_cost) {
        ORN_TRY(ensure_batch_cost_order(b, st));
        kp.order = b->order;
        kp.cost = b->cost;
    }
    LaunchArgs a{kp, b->d_items, b->d_work, b->acc, b->rng, b->fb, b->stage, b->queues, b->total_tiles};
    ORN_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    hipError_t e = hipMemcpyAsync(b->d_items, b->h_items, b->n * sizeof(BatchItem), hipMemcpyHostToDevice, st);
    if (e == hipSuccess && by_cost) {
        OrderShape o{};
        o.work = b->d_work;
        o.items = b->d_items;
        enqueue_order_kernels(o, b->total_tiles, b->cost, b->counts, b->totals, b->order, st);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        reset_queues_kernel<<<1, 64, 0, st>>>(b->queues);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = launch_render<true>(s->features, width, b->tuning.workgroups_per_cu,
                                static_cast<uint32_t>(s->ctx->compute_units), lds, st, ds, a);
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(st, &g);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    b->graph = g;
    if (e != hipSuccess) return from_hip(e);
    b->graph_ok = true;
    return ORN_OK;
}

orn_status batch_readback(orn_batch *b, size_t item, void *dst, const void *pool, size_t elem, size_t pixels)
{
    if (!b) return ORN_ERROR_INVALID_HANDLE;
    if (!dst || item >= b->n) return ORN_ERROR_INVALID_VALUE;
    const size_t n = size_t(b->width[item]) * b->height[item];
    if (pixels < n) return ORN_ERROR_SIZE_MISMATCH;
    ORN_TRY(hipSetDevice(b->scene->ctx->device));
    hipStream_t st = b->scene->ctx->stream;
    const char *src = static_cast<const char *>(pool) + size_t(b->px_base[item]) * elem;
    ORN_TRY(hipMemcpyAsync(dst, src, n * elem, hipMemcpyDeviceToHost, st));
    ORN_TRY(hipStreamSynchronize(st));
    return ORN_OK;
}

}  // namespace

extern "C" {

orn_status orn_batch_create(orn_ctx *ctx, const orn_batch_item *items, size_t n, orn_batch **out)
{
    if (!ctx) return ORN_ERROR_INVALID_HANDLE;
    if (!out || !items || n == 0 || n > 0xffffffu) return ORN_ERROR_INVALID_VALUE;
    *out = nullptr;
    uint64_t px = 0, tiles = 0;
    for (size_t i = 0; i < n; i++) {
        if (items[i].width == 0 || items[i].height == 0) return ORN_ERROR_INVALID_VALUE;
        px += uint64_t(items[i].width) * items[i].height;
        tiles += uint64_t(blocks_for(items[i].width, kTile)) * blocks_for(items[i].height, kTile);
    }
    if (px > kMaxPixels) return ORN_ERROR_INVALID_VALUE;
    MergedScene m;
    uint32_t slots = 0;
    orn_status st = merge_scenes(items, n, m, &slots);
    if (st != ORN_OK) return st;
    std::vector<uint32_t> root2(n), root4(n);
    orn_device_scene *s = nullptr;
    st = upload_scene(ctx, &m.desc, slots, m.roots.data(), n, root2.data(), root4.data(), &s);
    if (st != ORN_OK) return st;
    orn_batch *b = new (std::nothrow) orn_batch();
    if (!b) {
        release_scene(s);
        return ORN_ERROR_OUT_OF_MEMORY;
    }
    b->scene = s;  // the batch holds the scene's only reference
    b->n = n;
    b->root2 = root2;
    b->root4 = root4;
    b->tex_base = m.tex_base;
    b->tuning = default_tuning();
    uint32_t base = 0;
    for (size_t i = 0; i < n; i++) {
        b->width.push_back(items[i].width);
        b->height.push_back(items[i].height);
        b->tiles_x.push_back(blocks_for(items[i].width, kTile));
        b->tiles.push_back(b->tiles_x.back() * blocks_for(items[i].height, kTile));
        b->px_base.push_back(base);
        base += items[i].width * items[i].height;
    }
    b->total_px = base;
    b->total_tiles = static_cast<uint32_t>(tiles);
    b->work_bytes = sizeof(BatchTiles) + tiles * sizeof(uint2);
    auto fail = [&](hipError_t e) {
        free_batch(b);
        return from_hip(e);
    };
    hipError_t e;
    if ((e = hipHostMalloc(reinterpret_cast<void **>(&b->h_items), n * sizeof(BatchItem), hipHostMallocDefault)) !=
        hipSuccess)
        return fail(e);
    if ((e = hipHostMalloc(reinterpret_cast<void **>(&b->h_work), b->work_bytes, hipHostMallocDefault)) != hipSuccess)
        return fail(e);
    if ((e = hipMalloc(&b->d_items, n * sizeof(BatchItem))) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&b->d_work, b->work_bytes)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&b->fb, size_t(base) * sizeof(float4))) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&b->acc, size_t(base) * sizeof(float4))) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&b->rng, size_t(base) * sizeof(uint32_t))) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&b->queues, 64 * sizeof(uint32_t))) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&b->stage, size_t(base) * sizeof(uint32_t))) != hipSuccess) return fail(e);
    if ((e = hipEventCreate(&b->ev_start)) != hipSuccess) return fail(e);
    if ((e = hipEventCreate(&b->ev_stop)) != hipSuccess) return fail(e);
    std::memset(b->h_items, 0, n * sizeof(BatchItem));
    hipStream_t stream = ctx->stream;
    zero_kernel<<<blocks_for(base, 256), 256, 0, stream>>>(b->fb, b->acc, base);
    fill_u32_kernel<<<blocks_for(base, 256), 256, 0, stream>>>(b->stage, base, 0u);
    for (size_t i = 0; i < n; i++) {
        const uint32_t nn = items[i].width * items[i].height;
        seed_rng_kernel<<<blocks_for(nn, 256), 256, 0, stream>>>(b->rng + b->px_base[i], nn, items[i].seed * nn);
    }
    if ((e = hipGetLastError()) != hipSuccess) return fail(e);
    if ((e = hipStreamSynchronize(stream)) != hipSuccess) return fail(e);
    *out = b;
    return ORN_OK;
}

orn_status orn_batch_destroy(orn_batch *b)
{
    if (!b) return ORN_ERROR_INVALID_HANDLE;
    free_batch(b);
    return ORN_OK;
}

orn_status orn_batch_render_async(orn_batch *b, const orn_constant_params *params, const uint32_t *iterations)
{
    if (!b) return ORN_ERROR_INVALID_HANDLE;
    if (!params || !iterations) return ORN_ERROR_INVALID_VALUE;
    for (size_t i = 0; i < b->n; i++)
        if (params[i].width != b->width[i] || params[i].height != b->height[i]) return ORN_ERROR_SIZE_MISMATCH;
    orn_device_scene *s = b->scene;
    ORN_TRY(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    // the previous launch may still read the pinned staging buffers
    ORN_TRY(hipStreamSynchronize(st));
    const bool w4 = b->tuning.bvh_width == 4u;
    if (++b->epoch > 0xffffu) {  // tags wrap: clear the stage words
        fill_u32_kernel<<<blocks_for(b->total_px, 256), 256, 0, st>>>(b->stage, b->total_px, 0u);
        b->epoch = 1;
    }
    std::vector<uint64_t> cost(b->n + 1);
    uint32_t max_chunks = 1;
    for (size_t i = 0; i < b->n; i++) {
        BatchItem &it = b->h_items[i];
        it.p = params[i];
        it.iterations = iterations[i];
        it.tiles_x = b->tiles_x[i];
        it.px_base = b->px_base[i];
        it.root = w4 ? b->root4[i] : b->root2[i];
        it.tex_base = b->tex_base[i];
        plan_jobs(iterations[i], params[i].current_iteration, b->tuning.samples_per_job,
                  uint64_t(b->total_tiles) * kTile * kTile, &it.chunk, &it.chunks);
        it.epoch = b->epoch;
        max_chunks = std::max(max_chunks, it.chunks);
        cost[i] = uint64_t(std::max(iterations[i], 1u)) * (uint64_t(params[i].depth) + 1u);
    }
    cost[b->n] = max_chunks;
    if (cost != b->work_cost) {
        build_work(b, cost);
        ORN_TRY(hipMemcpyAsync(b->d_work, b->h_work, b->work_bytes, hipMemcpyHostToDevice, st));
    }
    if (!b->graph_ok) {
        const orn_status cs = capture_batch(b);
        if (cs != ORN_OK) return cs;
    }
    ORN_TRY(hipEventRecord(b->ev_start, st));
    ORN_TRY(hipGraphLaunch(b->exec, st));
    ORN_TRY(hipEventRecord(b->ev_stop, st));
    b->timed = true;
    return ORN_OK;
}

orn_status orn_batch_wait(orn_batch *b)
{
    if (!b) return ORN_ERROR_INVALID_HANDLE;
    ORN_TRY(hipSetDevice(b->scene->ctx->device));
    ORN_TRY(hipStreamSynchronize(b->scene->ctx->stream));
    return ORN_OK;
}

orn_status orn_batch_render(orn_batch *b, const orn_constant_params *params, const uint32_t *iterations)
{
    const orn_status s = orn_batch_render_async(b, params, iterations);
    if (s != ORN_OK) return s;
    return orn_batch_wait(b);
}

orn_status orn_batch_set_tuning(orn_batch *b, const orn_tuning *tuning)
{
    if (!b) return ORN_ERROR_INVALID_HANDLE;
    if (!tuning_valid(tuning)) return ORN_ERROR_INVALID_VALUE;
    const orn_tuning n = normalized(*tuning);
    if (std::memcmp(&n, &b->tuning, sizeof(orn_tuning)) != 0) {
        ORN_TRY(hipSetDevice(b->scene->ctx->device));
        ORN_TRY(hipStreamSynchronize(b->scene->ctx->stream));  // the graph may be running
        b->tuning = n;
        b->graph_ok = false;  // captured again at the next render
    }
    return ORN_OK;
}

orn_status orn_batch_get_tuning(orn_batch *b, orn_tuning *tuning)
{
    if (!b) return ORN_ERROR_INVALID_HANDLE;
    if (!tuning) return ORN_ERROR_INVALID_VALUE;
    *tuning = b->tuning;
    return ORN_OK;
}

orn_status orn_batch_last_render_ms(orn_batch *b, float *ms)
{
    if (!b) return ORN_ERROR_INVALID_HANDLE;
    if (!ms) return ORN_ERROR_INVALID_VALUE;
    *ms = 0.0f;
    if (!b->timed) return ORN_OK;
    ORN_TRY(hipEventSynchronize(b->ev_stop));
    ORN_TRY(hipEventElapsedTime(ms, b->ev_start, b->ev_stop));
    return ORN_OK;
}

orn_status orn_batch_get_framebuffer(orn_batch *b, size_t item, float *dst, size_t pixels)
{
    return batch_readback(b, item, dst, b ? b->fb : nullptr, sizeof(float4), pixels);
}

orn_status orn_batch_get_accumulation(orn_batch *b, size_t item, float *dst, size_t pixels)
{
    return batch_readback(b, item, dst, b ? b->acc : nullptr, sizeof(float4), pixels);
}

orn_status orn_batch_get_rng_state(orn_batch *b, size_t item, uint32_t *dst, size_t pixels)
{
    return batch_readback(b, item, dst, b ? b->rng : nullptr, sizeof(uint32_t), pixels);
}

orn_status orn_batch_set_state(orn_batch *b, size_t item, const float *accumulation, const uint32_t *rng,
                               size_t pixels)
{
    if (!b) return ORN_ERROR_INVALID_HANDLE;
    if (item >= b->n) return ORN_ERROR_INVALID_VALUE;
    const size_t n = size_t(b->width[item]) * b->height[item];
    if (pixels != n) return ORN_ERROR_SIZE_MISMATCH;
    ORN_TRY(hipSetDevice(b->scene->ctx->device));
    hipStream_t st = b->scene->ctx->stream;
    const size_t at = b->px_base[item];
    if (accumulation)
        ORN_TRY(hipMemcpyAsync(b->acc + at, accumulation, n * sizeof(float4), hipMemcpyHostToDevice, st));
    if (rng) ORN_TRY(hipMemcpyAsync(b->rng + at, rng, n * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    ORN_TRY(hipStreamSynchronize(st));
    return ORN_OK;
}

}  // extern "C"

// ===========================================================================
// Autotuning
// ===========================================================================

namespace {

// Flat JSON cache: {"<key>": {"<field>": <number>, ...}, ...}
using TuneCache = std::vector<std::pair<std::string, std::vector<std::pair<std::string, double>>>>;

void skip_ws(const std::string &t, size_t &i)
{
    while (i < t.size() && (t[i] == ' ' || t[i] == '\n' || t[i] == '\r' || t[i] == '\t')) i++;
}

bool parse_string(const std::string &t, size_t &i, std::string &out)
{
    skip_ws(t, i);
    if (i >= t.size() || t[i] != '"') return false;
    const size_t end = t.find('"', i + 1);
    if (end == std::string::npos) return false;
    out = t.substr(i + 1, end - i - 1);
    i = end + 1;
    return true;
}

bool expect(const std::string &t, size_t &i, char c)
{
    skip_ws(t, i);
    if (i >= t.size() || t[i] != c) return false;
    i++;
    return true;
}

// Reads the cache; a missing or malformed file reads as empty.
TuneCache load_cache(const char *path)
{
    TuneCache c;
    FILE *f = std::fopen(path, "rb");
    if (!f) return c;
    std::string t;
    char buf[4096];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) t.append(buf, n);
    std::fclose(f);
    size_t i = 0;
    if (!expect(t, i, '{')) return {};
    skip_ws(t, i);
    if (i < t.size() && t[i] == '}') return c;
    for (;;) {
        std::string key;
        if (!parse_string(t, i, key) || !expect(t, i, ':') || !expect(t, i, '{')) return {};
        std::vector<std::pair<std::string, double>> fields;
        skip_ws(t, i);
        if (i < t.size() && t[i] == '}') {
            i++;
        } else {
            for (;;) {
                std::string name;
                if (!parse_string(t, i, name) || !expect(t, i, ':')) return {};
                skip_ws(t, i);
                char *endp = nullptr;
                const double v = std::strtod(t.c_str() + i, &endp);
                if (endp == t.c_str() + i) return {};
                i = static_cast<size_t>(endp - t.c_str());
                fields.emplace_back(name, v);
                skip_ws(t, i);
                if (i < t.size() && t[i] == ',') {
                    i++;
                    continue;
                }
                if (!expect(t, i, '}')) return {};
                break;
            }
        }
        c.emplace_back(key, fields);
        skip_ws(t, i);
        if (i < t.size() && t[i] == ',') {
            i++;
            continue;
        }
        if (!expect(t, i, '}')) return {};
        return c;
    }
}

bool save_cache(const char *path, const TuneCache &c)
{
    const std::string tmp = std::string(path) + ".tmp";
    FILE *f = std::fopen(tmp.c_str(), "wb");
    if (!f) return false;
    std::fprintf(f, "{\n");
    for (size_t k = 0; k < c.size(); k++) {
        std::fprintf(f, "  \"%s\": {", c[k].first.c_str());
        for (size_t j = 0; j < c[k].second.size(); j++)
            std::fprintf(f, "%s\"%s\": %.9g", j ? ", " : "", c[k].second[j].first.c_str(), c[k].second[j].second);
        std::fprintf(f, "}%s\n", k + 1 < c.size() ? "," : "");
    }
    std::fprintf(f, "}\n");
    const bool ok = std::fclose(f) == 0;
    return ok && std::rename(tmp.c_str(), path) == 0;
}

uint32_t log2_bucket(uint64_t v)
{
    uint32_t b = 0;
    while (v) {
        b++;
        v >>= 1;
    }
    return b;
}

// Scene class: what decides the best launch shape, bucketed so that
// similar scenes share one entry.
std::string scene_class(const orn_device_scene *s, uint64_t pixels, uint32_t spp, uint32_t depth, size_t images,
                        uint32_t flags)
{
    char key[256];
    std::snprintf(key, sizeof(key), "%s|features%u|tlas2^%u|blas2^%u|px2^%u|spp2^%u|depth%u|images2^%u|bvh4%u",
                  s->ctx->arch, s->features, log2_bucket(s->ds.tlas_len), log2_bucket(s->blas_len),
                  log2_bucket(pixels), log2_bucket(std::max(spp, 1u)), depth, log2_bucket(images),
                  (flags & ORN_TUNE_ALLOW_BVH4) ? 1u : 0u);
    return key;
}

bool cached_tuning(const char *path, const std::string &key, orn_tuning *t, float *ms)
{
    if (!path) return false;
    const TuneCache c = load_cache(path);
    for (const auto &e : c) {
        if (e.first != key) continue;
        orn_tuning r{};
        bool ok[4] = {false, false, false, false};
        for (const auto &f : e.second) {
            const uint32_t v = static_cast<uint32_t>(f.second);
            if (f.first == "regroup_lanes") r.regroup_lanes = v, ok[0] = true;
            else if (f.first == "leaf_regroup_lanes") r.leaf_regroup_lanes = v, ok[1] = true;
            else if (f.first == "workgroups_per_cu") r.workgroups_per_cu = v, ok[2] = true;
            else if (f.first == "bvh_width") r.bvh_width = v, ok[3] = true;
            else if (f.first == "samples_per_job") r.samples_per_job = v;
            else if (f.first == "split_percent") r.split_percent = v;
            else if (f.first == "node_steps") r.node_steps = v;
            else if (f.first == "workgroup_pool") r.workgroup_pool = v;
            else if (f.first == "cost_order") r.cost_order = v;
            else if (f.first == "ms") *ms = static_cast<float>(f.second);
        }
        if (!(ok[0] && ok[1] && ok[2] && ok[3])) return false;
        if (!tuning_valid(&r)) return false;
        *t = normalized(r);
        return true;
    }
    return false;
}

void store_tuning(const char *path, const std::string &key, const orn_tuning &t, float ms)
{
    if (!path) return;
    TuneCache c = load_cache(path);
    std::vector<std::pair<std::string, double>> fields = {
        {"regroup_lanes", t.regroup_lanes},
        {"leaf_regroup_lanes", t.leaf_regroup_lanes},
        {"workgroups_per_cu", t.workgroups_per_cu},
        {"bvh_width", t.bvh_width},
        {"samples_per_job", t.samples_per_job},
        {"split_percent", t.split_percent},
        {"node_steps", t.node_steps},
        {"workgroup_pool", t.workgroup_pool},
        {"cost_order", t.cost_order},
        {"ms", ms}};
    bool found = false;
    for (auto &e : c)
        if (e.first == key) {
            e.second = fields;
            found = true;
        }
    if (!found) c.emplace_back(key, fields);
    (void)save_cache(path, c);
}

// Coordinate descent over the tuning fields; `run` renders once under a
// tuning and returns its GPU time. Each candidate keeps the faster of two
// renders.
template <class Run>
orn_status coordinate_descent(const orn_tuning &start, uint32_t flags, Run run, orn_tuning *best, float *best_ms)
{
    std::vector<std::pair<orn_tuning, float>> seen;
    auto measure = [&](const orn_tuning &t, float *ms) -> orn_status {
        for (const auto &e : seen)
            if (std::memcmp(&e.first, &t, sizeof(orn_tuning)) == 0) {
                *ms = e.second;
                return ORN_OK;
            }
        float a = 0, b = 0;
        orn_status st = run(t, &a);
        if (st == ORN_OK) st = run(t, &b);
        if (st != ORN_OK) return st;
        *ms = std::min(a, b);
        seen.emplace_back(t, *ms);
        return ORN_OK;
    };
    orn_tuning cur = start;
    float cur_ms = 0;
    orn_status st = measure(cur, &cur_ms);
    if (st != ORN_OK) return st;
    static const uint32_t regroup[] = {8, 16, 24, 32, 48, 64};
    static const uint32_t leaf[] = {1, 4, 8, 16, 32};
    static const uint32_t per_cu[] = {0, 1, 2, 3};
    static const uint32_t width[] = {2, 4};
    static const uint32_t split[] = {25, 50, 100};
    static const uint32_t job[] = {2, 4, 8, 16, 32};
    static const uint32_t steps[] = {1, 2, 4, 6, 8, 12};
    static const uint32_t off_on[] = {0, 1};
    auto sweep = [&](uint32_t orn_tuning::*field, const uint32_t *vals, size_t n) -> orn_status {
        for (size_t i = 0; i < n; i++) {
            orn_tuning t = cur;
            t.*field = vals[i];
            float ms = 0;
            const orn_status s = measure(t, &ms);
            if (s != ORN_OK) return s;
            if (ms < cur_ms) {
                cur = t;
                cur_ms = ms;
            }
        }
        return ORN_OK;
    };
    // Passes over every parameter until one changes nothing (at most 4): the
    // parameters interact (small jobs pay off only with wider shading rounds,
    // for instance), so one pass can stop far from the best setting.
    for (int pass = 0; pass < 4; pass++) {
        const float before = cur_ms;
        if ((st = sweep(&orn_tuning::regroup_lanes, regroup, 6)) != ORN_OK) return st;
        if ((st = sweep(&orn_tuning::leaf_regroup_lanes, leaf, 5)) != ORN_OK) return st;
        if ((st = sweep(&orn_tuning::node_steps, steps, 6)) != ORN_OK) return st;
        if ((st = sweep(&orn_tuning::regroup_lanes, regroup, 6)) != ORN_OK) return st;
        if ((st = sweep(&orn_tuning::workgroups_per_cu, per_cu, 4)) != ORN_OK) return st;
        // job splitting: both fields at once (they only pay off together)
        for (uint32_t sp : split)
            for (uint32_t jb : job) {
                orn_tuning t = cur;
                t.split_percent = sp;
                t.samples_per_job = jb;
                float ms = 0;
                if ((st = measure(t, &ms)) != ORN_OK) return st;
                if (ms < cur_ms) {
                    cur = t;
                    cur_ms = ms;
                }
            }
        if ((flags & ORN_TUNE_ALLOW_BVH4) && (st = sweep(&orn_tuning::bvh_width, width, 2)) != ORN_OK) return st;
        if ((st = sweep(&orn_tuning::cost_order, off_on, 2)) != ORN_OK) return st;
        if (!(cur_ms < before)) break;
    }
    // Then the settings between the coarse candidates around the best one:
    // regroup_lanes +-4, leaf_regroup_lanes +-2 and +-4, node_steps +-1.
    auto nudge = [&](uint32_t orn_tuning::*field, int delta, int lo, int hi) -> orn_status {
        const int v = static_cast<int>(cur.*field) + delta;
        if (v < lo || v > hi) return ORN_OK;
        orn_tuning t = cur;
        t.*field = static_cast<uint32_t>(v);
        float ms = 0;
        const orn_status s = measure(t, &ms);
        if (s != ORN_OK) return s;
        if (ms < cur_ms) {
            cur = t;
            cur_ms = ms;
        }
        return ORN_OK;
    };
    for (int d : {-4, 4})
        if ((st = nudge(&orn_tuning::regroup_lanes, d, 1, 64)) != ORN_OK) return st;
    for (int d : {-4, -2, 2, 4})
        if ((st = nudge(&orn_tuning::leaf_regroup_lanes, d, 1, 64)) != ORN_OK) return st;
    for (int d : {-1, 1})
        if ((st = nudge(&orn_tuning::node_steps, d, 1, 16)) != ORN_OK) return st;
    *best = cur;
    *best_ms = cur_ms;
    return ORN_OK;
}

// Device copies of a target's or batch's pixel state, restored after
// every trial render.
struct SavedState {
    void *acc = nullptr, *rng = nullptr, *fb = nullptr;
    size_t px = 0;
    hipStream_t st = nullptr;
    ~SavedState()
    {
        if (st) (void)hipStreamSynchronize(st);
        if (acc) (void)hipFree(acc);
        if (rng) (void)hipFree(rng);
        if (fb) (void)hipFree(fb);
    }
    hipError_t save(hipStream_t s, const float4 *a, const uint32_t *r, const float4 *f, size_t n)
    {
        st = s;
        px = n;
        hipError_t e = hipMalloc(&acc, n * sizeof(float4));
        if (e == hipSuccess) e = hipMalloc(&rng, n * sizeof(uint32_t));
        if (e == hipSuccess) e = hipMalloc(&fb, n * sizeof(float4));
        if (e == hipSuccess) e = hipMemcpyAsync(acc, a, n * sizeof(float4), hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(rng, r, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(fb, f, n * sizeof(float4), hipMemcpyDeviceToDevice, s);
        return e;
    }
    hipError_t restore(float4 *a, uint32_t *r, float4 *f) const
    {
        hipError_t e = hipMemcpyAsync(a, acc, px * sizeof(float4), hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(r, rng, px * sizeof(uint32_t), hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(f, fb, px * sizeof(float4), hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        return e;
    }
};

}  // namespace

extern "C" {

orn_status orn_target_autotune(orn_target *t, const orn_constant_params *params, uint32_t iterations, uint32_t flags,
                               const char *cache_path, orn_tuning *chosen, float *ms)
{
    if (!t) return ORN_ERROR_INVALID_HANDLE;
    if (!params) return ORN_ERROR_INVALID_VALUE;
    if (params->width != t->width || params->height != t->height) return ORN_ERROR_SIZE_MISMATCH;
    orn_device_scene *s = t->scene;
    const std::string key =
        scene_class(s, uint64_t(t->width) * t->height, iterations, params->depth, 1, flags);
    orn_tuning best{};
    float best_ms = 0.0f;
    if (!(flags & ORN_TUNE_IGNORE_CACHE) && cached_tuning(cache_path, key, &best, &best_ms)) {
        t->tuning = best;
    } else {
        ORN_TRY(hipSetDevice(s->ctx->device));
        hipStream_t st = s->ctx->stream;
        ORN_TRY(hipStreamSynchronize(st));
        SavedState saved;
        ORN_TRY(saved.save(st, t->acc, t->rng, t->fb, size_t(t->width) * t->height));
        const orn_tuning start = t->tuning;
        auto run = [&](const orn_tuning &tn, float *m) -> orn_status {
            t->tuning = tn;
            orn_status r = orn_render(t, params, iterations);
            if (r == ORN_OK) r = orn_target_last_render_ms(t, m);
            const hipError_t e = saved.restore(t->acc, t->rng, t->fb);
            return r != ORN_OK ? r : from_hip(e);
        };
        const orn_status r = coordinate_descent(start, flags, run, &best, &best_ms);
        t->tuning = r == ORN_OK ? best : start;
        if (r != ORN_OK) return r;
        store_tuning(cache_path, key, best, best_ms);
    }
    if (chosen) *chosen = best;
    if (ms) *ms = best_ms;
    return ORN_OK;
}

orn_status orn_batch_autotune(orn_batch *b, const orn_constant_params *params, const uint32_t *iterations,
                              uint32_t flags, const char *cache_path, orn_tuning *chosen, float *ms)
{
    if (!b) return ORN_ERROR_INVALID_HANDLE;
    if (!params || !iterations) return ORN_ERROR_INVALID_VALUE;
    uint32_t spp = 0, depth = 0;
    for (size_t i = 0; i < b->n; i++) {
        if (params[i].width != b->width[i] || params[i].height != b->height[i]) return ORN_ERROR_SIZE_MISMATCH;
        spp = std::max(spp, iterations[i]);
        depth = std::max(depth, params[i].depth);
    }
    orn_device_scene *s = b->scene;
    const std::string key = scene_class(s, b->total_px, spp, depth, b->n, flags);
    orn_tuning best{};
    float best_ms = 0.0f;
    if (!(flags & ORN_TUNE_IGNORE_CACHE) && cached_tuning(cache_path, key, &best, &best_ms)) {
        const orn_status r = orn_batch_set_tuning(b, &best);
        if (r != ORN_OK) return r;
    } else {
        ORN_TRY(hipSetDevice(s->ctx->device));
        hipStream_t st = s->ctx->stream;
        ORN_TRY(hipStreamSynchronize(st));
        SavedState saved;
        ORN_TRY(saved.save(st, b->acc, b->rng, b->fb, b->total_px));
        const orn_tuning start = b->tuning;
        auto run = [&](const orn_tuning &tn, float *m) -> orn_status {
            orn_status r = orn_batch_set_tuning(b, &tn);
            if (r == ORN_OK) r = orn_batch_render(b, params, iterations);
            if (r == ORN_OK) r = orn_batch_last_render_ms(b, m);
            const hipError_t e = saved.restore(b->acc, b->rng, b->fb);
            return r != ORN_OK ? r : from_hip(e);
        };
        const orn_status r = coordinate_descent(start, flags, run, &best, &best_ms);
        (void)orn_batch_set_tuning(b, r == ORN_OK ? &best : &start);
        if (r != ORN_OK) return r;
        store_tuning(cache_path, key, best, best_ms);
    }
    if (chosen) *chosen = best;
    if (ms) *ms = best_ms;
    return ORN_OK;
}

}  // extern "C"
