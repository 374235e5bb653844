"""Benchmark of the path-tracing hot path on one MI355X.

Workload: the random-spheres scene (C2: ground + 22x22 jittered grid of
r=0.2 spheres + three r=1 spheres), 1920x1080, 64 samples per pixel, depth 16,
seed 0. One step = one PathTracer.render() of 64 spp over the whole image
(wgpu_backend/path_tracer.zig:252-265), i.e. 132,710,400 samples, with the
scene already resident in HBM. Metric: samples/s = W*H*spp*steps / wall time.
`--workload C1` (Cornell box 256x256, 16 spp, depth 8) or `--workload C3` (1,024
instanced icospheres, 1920x1080, 32 spp, depth 8) times those scenes the same
way (the profiled HBM traffic is reported for C2 only).

Launch tuning: the library's autotuner picks the render kernel's launch
parameters for the scene class; its JSON cache ships with the package
(zig_ornament_amd/tuning_cache.json), so a known class is applied without
trial renders. Tuning never changes the image. C2's tuning queues each XCD's
pixels by what they cost in the target's previous render (cost_order: most
expensive first, so the lanes run out of work nearly together); the warm-up
renders record those costs, and `first_render_kernel_ms` reports the
target's first render, which has none yet and records them in its first
eighth of the samples (two launches).

Also reported:
  roofline     the bound that limits the render kernel: vector-instruction
               issue. `achieved` = the lanes that did VALU work per launch
               (SQ_THREAD_CYCLES_VALU, from the committed rocprofv3 counter
               summary of this workload and launch) divided by the kernel's
               average duration (HIP events on the stream it runs on);
               `peak` = 39.3e12 lane-operations/s, MI355X's published FP32
               vector peak (157.3 TFLOP/s) without its factor 2 for FMA and 2
               for packed FP32 (1,024 SIMDs x 16 lanes x 2.4 GHz). Also:
               `lane_occupancy` (active lanes per VALU instruction / 64),
               `valu_issue_frac` (VALU issue share of SIMD cycles), and `hbm`:
               HBM bytes per launch from FETCH_SIZE/WRITE_SIZE (scaled by a
               plain copy kernel) over the live kernel time against 8 TB/s,
               beside the reference algorithm's bytes per second (node,
               transform, material and attribute fetches counted on the same
               paths by the instrumented CPU oracle), which on-chip caches
               serve and which therefore exceed the HBM peak.
  cpu_baseline the oracle (a C restatement of the WGSL megakernel, OpenMP
               over pixels, serial samples per pixel) timed on the host on a
               bounded sample of the same workload, on every CPU the process
               may use up to the box's allotment for one GPU (16 threads);
               nproc, os.cpu_count() and the CPU model are reported.
  workloads    the other configurations timed the same way (median kernel
               time of three renders after a warm-up, shipped tuning): the
               Cornell box (C1), the instanced icospheres (C3), and the batch
               of 64 instanced-icosphere scenes in one graph-replayed launch
               (C4).
  equal_time_rmse
               plain path tracing vs the neural radiance cache on the
               instanced-icosphere scene (1,024 instances of a 327,680-triangle
               icosphere, two 1024^2 textures, 1920x1080, depth 8): the cache
               trains online along a seeded 120-frame camera path at 1 spp per
               frame; the last frame's 1-spp image is compared (RMSE and
               relMSE on linear radiance, and RMSE of the post-processed
               framebuffers) with a 4,096-spp plain render of that
               view, next to plain renders of the same view at increasing spp:
               the plain RMSE in the cache's frame time, and the plain time
               needed to reach the cache's RMSE. Two training schedules:
               `nrc` as the workload specifies it (2^19-entry tables, 65,536
               training paths per frame, one Adam step over their records)
               and `nrc_light` (16,384 paths, one step).
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import zig_ornament_amd as zo  # noqa: E402
from zig_ornament_amd import workloads as W  # noqa: E402

HBM_PEAK_GBS = 8000.0
# --workload: (config name, metric text); C2 is the headline configuration
WORKLOADS = {
    "C1": ("cornell_256x256_16spp_depth8", "Cornell box 256x256 16spp depth 8"),
    "C2": ("random_spheres_1920x1080_64spp_depth16", "random-spheres 1920x1080 64spp depth 16"),
    "C3": ("instanced_icospheres_1920x1080_32spp_depth8", "1,024 instanced icospheres 1920x1080 32spp depth 8"),
}
# committed rocprofv3 counter summaries of the render kernel per workload:
# HBM traffic (FETCH_SIZE / WRITE_SIZE) and instruction counters (SQ_*)
PROFILES = {
    "C2": ("render_kernel_random_spheres_1080p_64spp_pmc.md", "render_kernel_random_spheres_1080p_64spp_sq.md"),
    "C3": ("render_kernel_instanced_icospheres_1080p_32spp_pmc.md",
           "render_kernel_instanced_icospheres_1080p_32spp_sq.md"),
}
# vector lanes the SIMDs can drive per second: 157.3 TFLOP/s FP32 (FMA = 2,
# packed FP32 = 2) / 4 = 1,024 SIMDs x 16 lanes x 2.4 GHz
VALU_PEAK_LANE_OPS = 157.3e12 / 4


def profiled_figure(path, name, kind=float):
    """A figure (`name = N`) from a committed rocprofv3 counter summary, or
    None."""
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(name + r"\s*=\s*([0-9.]+)", text)
    return kind(m.group(1)) if m else None


def profile_path(workload, kind):
    names = PROFILES.get(workload)
    return os.path.join(ROOT, "profiles", names[kind]) if names else ""


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def host_threads():
    """The CPUs this process may run on, at most the box's allotment for one
    GPU (OMP_NUM_THREADS, 16 by default)."""
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    return max(1, min(avail, cap)), avail


def algorithmic_bytes(stats):
    """Bytes the reference algorithm fetches (per the counters of one run):
    64 B per BVH node and per transform fetched during traversal; per shaded
    hit 32 B of material, +64 B model transform for spheres, +160 B of
    indices/normals/uvs/normal transform for meshes, +4 B per texel; 40 B per
    sample of RNG state and accumulation read+write."""
    hits = stats["sphere_hits"] + stats["mesh_hits"]
    return (64 * stats["node_visits"] + 64 * stats["transform_fetches"] + 32 * hits
            + 64 * stats["sphere_hits"] + (12 + 48 + 12 + 24 + 64) * stats["mesh_hits"]
            + 4 * stats["textured_hits"] + 40 * stats["samples"])


def cpu_baseline(spec, arrays, params, seconds=12.0):
    """Times the oracle on a bounded, uniformly spread subset of the pixels at
    full spp; returns samples/s, the sample size, threads and per-sample
    algorithmic bytes of those paths."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    O.build()
    threads, _ = host_threads()
    n_px = spec.width * spec.height
    sc = O.OracleScene(arrays)
    p = O.ConstantParams.from_buffer_copy(bytes(params))
    count = 2048
    while True:
        pix = np.linspace(0, n_px - 1, count).astype(np.uint32)
        acc = np.zeros((count, 4), np.float32)
        rng = pix.copy()
        st = O.Stats()
        t0 = time.perf_counter()
        sc.render_pixels(p, spec.spp, pix, acc, rng, nthreads=threads, stats=st)
        dt = time.perf_counter() - t0
        if dt >= seconds or count >= n_px:
            break
        count = min(n_px, max(count * 2, int(count * seconds / max(dt, 1e-3) * 1.05)))
    stats = {f: getattr(st, f) for f, _ in O.Stats._fields_}
    return dict(samples_per_s=count * spec.spp / dt, pixels=count, seconds=dt, threads=threads,
                stats=stats, bytes_per_sample=algorithmic_bytes(stats) / stats["samples"])


def rmse(img, ref):
    """RMSE over RGB of linear radiance; pixels with a non-finite sample in
    either image (the reference estimator's rare NaN at the total-internal-
    reflection boundary) are left out of both."""
    d = np.asarray(img, np.float64)[:, :3] - np.asarray(ref, np.float64)[:, :3]
    ok = np.all(np.isfinite(d), axis=1)
    return float(np.sqrt(np.mean(d[ok] ** 2)))


def relmse(img, ref):
    """mean((x - r)^2 / (r^2 + 0.01)) over RGB of linear radiance, with the
    same masking of non-finite pixels as rmse()."""
    x = np.asarray(img, np.float64)[:, :3]
    r = np.asarray(ref, np.float64)[:, :3]
    ok = np.all(np.isfinite(x - r), axis=1)
    return float(np.mean((x[ok] - r[ok]) ** 2 / (r[ok] ** 2 + 0.01)))


# The radiance-cache runs of the equal-time comparison: the cache as the
# workload names it (2^19-entry tables, 65,536 training paths per frame,
# their records taken by one Adam step), and a lighter schedule (16,384
# paths, one step) that spends less of each frame on training.
NRC_RUNS = (
    ("nrc", dict(log2_table=19, train_paths=65536, train_batch=65536), 1),
    ("nrc_light", dict(log2_table=19, train_paths=16384, train_batch=16384), 1),
)


def equal_time_rmse(dev, frames=120, ref_spp=4096):
    """The cache against plain path tracing at equal GPU time (see the module
    docstring)."""
    from zig_ornament_amd.nrc import Nrc
    spec = W.instanced_icospheres(spp=1)  # 1920x1080, depth 8
    scene = W.build_scene(spec)
    ds = zo.DeviceScene(dev, zo.Bvh(scene))
    lookat, vup = spec.camera[1], spec.camera[2]

    def params_at(frame):
        scene.set_look_at(W.orbit_lookfrom(spec, frame, frames), lookat, vup)
        p = zo._native.ConstantParams()
        p.camera = scene.camera()
        p.depth, p.width, p.height = spec.depth, spec.width, spec.height
        p.flip_y, p.inverted_gamma = 0, 1.0
        p.ray_cast_epsilon = spec.ray_cast_epsilon
        p.textures_count = len(spec.textures)
        p.current_iteration = 0.0
        return p

    n = spec.width * spec.height
    p = params_at(frames - 1)
    ref = zo.Target(ds, spec.width, spec.height, 99)
    ref.autotune(p, 32, cache_path=zo.TUNING_CACHE)
    done = 0
    while done < ref_spp:  # progressive renders of at most 512 spp each
        k = min(512, ref_spp - done)
        ref.render(_with_iteration(p, done), k)
        done += k
    acc = ref.accumulation()
    reference = acc / acc[:, 3:4]
    reference_fb = ref.framebuffer()  # post-processed (pathtracer.wgsl:81-93)

    runs = {}
    for name, cfg, steps in NRC_RUNS:
        nrc = Nrc(ds, terminate_bounce=0, **cfg)
        frame_ms = []
        t = zo.Target(ds, spec.width, spec.height, 7)
        for f in range(frames):  # each frame: a new view; 1-spp render beside training paths, then the steps
            ms, _ = nrc.frame(t, params_at(f), 1, f, steps=steps)
            frame_ms.append(ms)
        img = t.accumulation()
        runs[name] = {"rmse": rmse(img, reference), "relmse": relmse(img, reference),
                      "rmse_framebuffer": rmse(t.framebuffer(), reference_fb),
                      "ms_per_frame": float(np.median(frame_ms[frames // 2:])), "spp": 1,
                      "query_bounce": 0, "log2_table": cfg["log2_table"],
                      "train_paths_per_frame": cfg["train_paths"], "adam_steps_per_frame": steps,
                      "batch": cfg["train_batch"]}
        nrc.destroy()

    target = min(r["rmse"] for r in runs.values())
    slowest = max(r["ms_per_frame"] for r in runs.values())
    plain = []
    pt = zo.Target(ds, spec.width, spec.height, 7)
    for spp in (1, 2, 3, 4, 6, 8, 12, 16, 24, 32):
        pt.autotune(p, spp, cache_path=zo.TUNING_CACHE)  # the plain tracer's best launch for this spp
        ms = []
        for _ in range(3):  # median of three renders of the same samples
            pt.set_state(np.zeros((n, 4), np.float32), np.arange(n, dtype=np.uint32) + np.uint32(7 * n))
            pt.render(p, spp)
            ms.append(pt.last_render_ms())
        img = pt.accumulation() / spp
        plain.append({"spp": spp, "ms": float(np.median(ms)), "rmse": rmse(img, reference),
                      "relmse": relmse(img, reference), "rmse_framebuffer": rmse(pt.framebuffer(), reference_fb)})
        if spp >= 8 and plain[-1]["rmse"] <= target and plain[-1]["ms"] > slowest:
            break
    out = {"scene": "instanced_icospheres_1920x1080_depth8", "frames": frames, "reference_spp": ref_spp}
    for name, r in runs.items():
        within = [q for q in plain if q["ms"] <= r["ms_per_frame"]]
        match = next((q for q in plain if q["rmse"] <= r["rmse"]), None)
        r["plain_at_equal_time"] = within[-1] if within else None
        if not within:
            r["plain_at_equal_time_note"] = (f"one plain sample per pixel takes {plain[0]['ms']:.2f} ms, "
                                             f"longer than the cache's frame")
        r["plain_to_match_rmse"] = match
        r["plain_time_over_nrc_time"] = match["ms"] / r["ms_per_frame"] if match else None
        fit = equal_error_plain(plain, r["rmse"])
        r["plain_equal_error_interpolated"] = fit
        r["plain_time_over_nrc_time_interpolated"] = fit["ms"] / r["ms_per_frame"] if fit else None
        out[name] = r
    out["plain_curve"] = plain
    return out


def equal_error_plain(plain, target, max_spp=8):
    """The plain tracer's samples per pixel and time at exactly the RMSE
    `target`, interpolated instead of rounded up to a whole sample count:
    MSE(spp) = a + b / spp (independent samples: variance falls as 1/spp; a
    holds the reference's own noise) and time(spp) = c + d * spp, both fitted
    by least squares on the measured points up to `max_spp`."""
    pts = [q for q in plain if q["spp"] <= max_spp]
    if len(pts) < 3:
        return None
    spp = np.array([q["spp"] for q in pts], np.float64)
    mse = np.array([q["rmse"] for q in pts], np.float64) ** 2
    ms = np.array([q["ms"] for q in pts], np.float64)
    (a, b), *_ = np.linalg.lstsq(np.stack([np.ones_like(spp), 1 / spp], 1), mse, rcond=None)
    (c, d), *_ = np.linalg.lstsq(np.stack([np.ones_like(spp), spp], 1), ms, rcond=None)
    if target ** 2 <= a or b <= 0:
        return None
    s_star = b / (target ** 2 - a)
    return {"spp": float(s_star), "ms": float(c + d * s_star), "fit_points": [int(x) for x in spp],
            "mse_fit": [float(a), float(b)], "ms_fit": [float(c), float(d)]}


def time_target_workload(dev, key, steps=3):
    """Median render-kernel time of one of the single-image configurations
    with its shipped tuning (three renders after a warm-up)."""
    spec = W.CONFIGS[key]()
    ds = zo.DeviceScene(dev, zo.Bvh(W.build_scene(spec)))
    t = zo.Target(ds, spec.width, spec.height, 0)
    params = W.spec_params(spec)
    t.autotune(params, spec.spp, cache_path=zo.TUNING_CACHE)
    t.render(params, spec.spp)
    ms, wall = [], []
    for _ in range(steps):
        t0 = time.perf_counter()
        t.render(params, spec.spp)
        wall.append((time.perf_counter() - t0) * 1e3)
        ms.append(t.last_render_ms())
    med = float(np.median(ms))
    out = {"workload": WORKLOADS[key][0], "samples": spec.samples, "kernel_ms_median": med,
           "wall_ms_median": float(np.median(wall)), "samples_per_s": spec.samples / (med * 1e-3),
           "launch_tuning": t.tuning()}
    t.destroy()
    ds.destroy()
    return out


def time_batch_workload(dev, steps=3):
    """The same for C4: 64 instanced-icosphere scenes in one graph-replayed
    persistent launch (workloads.batch_scenes)."""
    specs = W.batch_scenes()
    arrays = [zo.Bvh(W.build_scene(s)) for s in specs]
    batch = zo.Batch(dev, [(a, s.width, s.height, s.seed) for a, s in zip(arrays, specs)])
    params = [W.spec_params(s) for s in specs]
    spp = [s.spp for s in specs]
    batch.autotune(params, spp, cache_path=zo.TUNING_CACHE)
    batch.render(params, spp)
    ms, wall = [], []
    for _ in range(steps):
        t0 = time.perf_counter()
        batch.render(params, spp)
        wall.append((time.perf_counter() - t0) * 1e3)
        ms.append(batch.last_render_ms())
    samples = sum(s.samples for s in specs)
    med = float(np.median(ms))
    out = {"workload": "batch_64_instanced_icosphere_scenes_128-1024px_1-64spp_depth8", "images": len(specs),
           "samples": samples, "kernel_ms_median": med, "wall_ms_median": float(np.median(wall)),
           "samples_per_s": samples / (med * 1e-3), "launch_tuning": batch.tuning()}
    batch.destroy()
    return out


def _with_iteration(p, it):
    q = zo._native.ConstantParams.from_buffer_copy(bytes(p))
    q.current_iteration = float(it)
    return q


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--calibration-copy-mib", type=int, default=0,
                    help="also run one plain copy kernel of this size (for scaling profiler "
                         "traffic counters); reported as calibration_copy")
    ap.add_argument("--tuning-cache", default=zo.TUNING_CACHE,
                    help="autotuner cache (JSON); a scene class missing from it is measured")
    ap.add_argument("--no-autotune", action="store_true", help="use the default launch tuning")
    ap.add_argument("--no-nrc", action="store_true", help="skip the equal-time RMSE leg")
    ap.add_argument("--no-workloads", action="store_true",
                    help="skip timing the other configurations (C1, C3, C4)")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="C2",
                    help="the timed scene (C2, random spheres, is the headline; C1 and C3 print the "
                         "same line for the Cornell box and the instanced icospheres)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the path tracer would shard an image by pixel tiles, "
              "with no communication while rendering and one gather of the framebuffer at the "
              "end, which is not implemented; it runs on a single MI355X. Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    spec = W.CONFIGS[args.workload]()
    workload, description = WORKLOADS[args.workload]
    t0 = time.perf_counter()
    scene = W.build_scene(spec)
    arrays = zo.Bvh(scene)
    build_s = time.perf_counter() - t0
    dev = zo.Device(0)
    name, cus = dev.info()
    ds = zo.DeviceScene(dev, arrays)
    tgt = zo.Target(ds, spec.width, spec.height, 0)
    params = zo._native.ConstantParams()
    params.camera = scene.camera()
    params.depth, params.width, params.height = spec.depth, spec.width, spec.height
    params.flip_y, params.inverted_gamma = 0, 1.0
    params.ray_cast_epsilon = spec.ray_cast_epsilon
    params.textures_count = len(spec.textures)
    params.current_iteration = 0.0

    tuning_ms = None
    if not args.no_autotune:
        _, tuning_ms = tgt.autotune(params, spec.spp, cache_path=args.tuning_cache)
    tuning = tgt.tuning()

    calib = None
    if args.calibration_copy_mib:
        nbytes = args.calibration_copy_mib << 20
        ms = dev.calibration_copy(nbytes)
        calib = {"bytes_read": nbytes, "bytes_written": nbytes, "ms": ms,
                 "GB_per_s": 2 * nbytes / (ms * 1e-3) / 1e9}
    first_ms = None
    for i in range(args.warmup):
        tgt.render(params, spec.spp)
        if i == 0:
            first_ms = tgt.last_render_ms()  # no previous costs yet
    tgt.wait()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        tgt.render(params, spec.spp)  # blocks until the GPU is done
        kernel_ms.append(tgt.last_render_ms())
    tgt.wait()
    wall = time.perf_counter() - t0
    samples = spec.samples * args.steps
    value = samples / wall
    avg_kernel_s = float(np.mean(kernel_ms)) / 1e3

    out = {
        "metric": f"samples/s (plain path tracing, {description}); "
                  "equal-time RMSE vs 4096-spp reference, plain PT vs NRC",
        "value": value,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded SplitMix64 scene generator)",
        "config": {"workload": workload, "width": spec.width,
                   "height": spec.height, "spp": spec.spp, "depth": spec.depth,
                   "spheres": len(spec.spheres), "tlas_nodes": int(len(arrays.tlas_nodes)),
                   "parallelism": "single GPU, pixel-parallel", "device": name,
                   "compute_units": cus, "bvh_build_s": build_s, "launch_tuning": tuning,
                   "autotuner_ms": tuning_ms},
        "kernel_ms_avg": avg_kernel_s * 1e3,
        "first_render_kernel_ms": first_ms,
    }
    if calib:
        out["calibration_copy"] = calib
    # the bound: vector issue at partial lane occupancy (counters from the
    # committed summary of this workload's launch, time measured live)
    sq_path, pmc_path = profile_path(args.workload, 1), profile_path(args.workload, 0)
    lane_ops = profiled_figure(sq_path, "valu_lane_ops_per_launch")
    traffic = profiled_figure(pmc_path, "traffic_bytes_per_launch", int)
    roof = {"bound": "valu", "achieved": None, "peak": VALU_PEAK_LANE_OPS / 1e12, "unit": "Tlane-op/s",
            "frac": None, "traffic": traffic,
            "lane_occupancy": profiled_figure(sq_path, "lane_occupancy"),
            "valu_issue_frac": profiled_figure(sq_path, "valu_issue_fraction"),
            "counters": os.path.relpath(sq_path, ROOT) if lane_ops else None}
    if lane_ops:
        roof["achieved"] = lane_ops / avg_kernel_s / 1e12
        roof["frac"] = roof["achieved"] / roof["peak"]
    hbm = {"peak_GBps": HBM_PEAK_GBS}
    if traffic:
        hbm["measured_GBps"] = traffic / avg_kernel_s / 1e9
        hbm["frac"] = hbm["measured_GBps"] / HBM_PEAK_GBS
    roof["hbm"] = hbm
    out["roofline"] = roof
    if not args.no_cpu_baseline:
        cb = cpu_baseline(spec, arrays, params, args.cpu_seconds)
        bytes_per_launch = cb["bytes_per_sample"] * spec.samples + 16 * spec.width * spec.height
        hbm["algorithmic_GBps"] = bytes_per_launch / avg_kernel_s / 1e9
        roof.update({"algorithmic_bytes_per_sample": cb["bytes_per_sample"],
                     "rays_per_sample": cb["stats"]["rays"] / cb["stats"]["samples"],
                     "node_visits_per_ray": cb["stats"]["node_visits"] / cb["stats"]["rays"],
                     "rays_per_s": value * cb["stats"]["rays"] / cb["stats"]["samples"]})
        threads, avail = host_threads()
        out["cpu_baseline"] = {"value": cb["samples_per_s"], "unit": "samples/s",
                               "cores": cb["threads"], "kind": "port",
                               "sample": f"{cb['pixels']} evenly spaced pixels of the same image "
                                         f"at {spec.spp} spp ({cb['pixels'] * spec.spp} samples, "
                                         f"{cb['seconds']:.1f} s) on {cb['threads']} host threads",
                               "nproc": avail, "os_cpu_count": os.cpu_count(), "cpu_model": cpu_model()}
        out["gpu_over_cpu"] = value / cb["samples_per_s"]
    if not args.no_workloads and args.workload == "C2":
        out["workloads"] = {k: time_target_workload(dev, k) for k in ("C1", "C3")}
        out["workloads"]["C4"] = time_batch_workload(dev)
    if not args.no_nrc:
        out["equal_time_rmse"] = equal_time_rmse(dev)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
