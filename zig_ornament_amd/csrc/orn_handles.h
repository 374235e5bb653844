// orn_handles.h -- the host-side objects behind the opaque handles of
// include/ornament_hip.h, shared by the translation units of the library.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/ornament_hip.h"
#include "orn_trace.h"

// Handles are reference counted: a scene holds its context and a target its
// scene, so the caller may destroy them in any order (a garbage-collected
// host language does).
struct orn_ctx {
    int device;
    hipStream_t stream;
    char name[256];
    char arch[64];  // gcnArchName, the autotuner's cache key
    int compute_units;
    int refs;
};

struct orn_device_scene {
    orn_ctx *ctx;
    orn::DevScene ds;
    std::vector<void *> allocations;
    uint32_t stack_slots;   // BVH2 traversal
    uint32_t stack_slots4;  // BVH4 traversal
    uint32_t trace_width;   // node width orn_trace_rays uses
    uint32_t features;      // SceneFeatures present in the arrays
    uint32_t blas_len;      // BVH2 BLAS nodes
    float bounds_lo[3];     // box of the whole scene (the TLAS root's)
    float bounds_hi[3];
    int refs;
};

struct orn_target {
    orn_device_scene *scene;
    orn_tuning tuning;
    uint32_t width, height;
    float4 *fb;
    float4 *acc;
    uint32_t *rng;
    uint32_t *stage;   // per pixel: launch tag << 16 | jobs finished in that launch
    uint32_t epoch;    // tag of the latest launch
    uint32_t *queues;  // one job-queue counter per XCD band
    uint32_t grid;     // persistent workgroups of the render kernel
    hipEvent_t ev_start, ev_stop;
    bool timed;
    // Cost-ordered queues (tuning cost_order), allocated by the first render
    // that asks for them.
    uint32_t *cost = nullptr;    // per pixel: GPU time of its jobs in the recording render
    uint32_t *order = nullptr;   // per band position: the band's tile-pixel queued there
    uint32_t *counts = nullptr;  // per tile and cost class: pixels, then their prefix over the band
    uint32_t *totals = nullptr;  // per band and cost class: pixels
    bool costs_recorded = false; // `cost` holds a whole render's recording
};

namespace orn_rt {

inline void release_ctx(orn_ctx *c)
{
    if (--c->refs > 0) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

inline void release_scene(orn_device_scene *s)
{
    if (--s->refs > 0) return;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    for (void *p : s->allocations) (void)hipFree(p);
    orn_ctx *c = s->ctx;
    delete s;
    release_ctx(c);
}

inline orn_status from_hip(hipError_t e) { return e == hipSuccess ? ORN_OK : static_cast<orn_status>(e); }

inline uint32_t blocks_for(uint32_t n, uint32_t b) { return (n + b - 1) / b; }

}  // namespace orn_rt

#define ORN_TRY(expr)                                       \
    do {                                                    \
        hipError_t e_ = (expr);                             \
        if (e_ != hipSuccess) return orn_rt::from_hip(e_);  \
    } while (0)
