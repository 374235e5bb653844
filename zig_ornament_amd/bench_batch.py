"""Benchmark of a batched render: 64 scenes in one persistent kernel launch
replayed from a captured HIP graph (orn_batch_render).

    python -m zig_ornament_amd.bench_batch [--steps K] [--warmup W] [--sequential]

Workload: workloads.batch_scenes() -- 64 instanced-icosphere scenes with
triangle counts log-uniform in [1e3, 2e6], square resolutions 128..1024 and
1..64 samples per pixel, depth 8, each image with its own RNG seed. One step =
one orn_batch_render of every image (all of its samples), with the scenes
resident in HBM. The launch tuning comes from the shipped tuning cache
(zig_ornament_amd/tuning_cache.json; a class missing from it is measured).
Prints one JSON line: samples/s of the batch, wall ms per batch, the render
kernel's own time (HIP events on its stream) and, with --sequential, the
same 64 renders issued one target after another (each with its own cached
tuning) for comparison.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import TUNING_CACHE, Batch, Bvh, Device, DeviceScene, Target
from . import workloads as W


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--sequential", action="store_true", help="also time the 64 renders one by one")
    ap.add_argument("--device", type=int, default=0)
    args = ap.parse_args(argv)
    specs = W.batch_scenes()
    t0 = time.perf_counter()
    arrays = [Bvh(W.build_scene(s)) for s in specs]
    build_s = time.perf_counter() - t0
    dev = Device(args.device)
    batch = Batch(dev, [(a, s.width, s.height, s.seed) for a, s in zip(arrays, specs)])
    params = [W.spec_params(s) for s in specs]
    spp = [s.spp for s in specs]
    tuning, tune_ms = batch.autotune(params, spp, cache_path=TUNING_CACHE)
    for _ in range(args.warmup):
        batch.render(params, spp)
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.render(params, spp)
        kernel_ms.append(batch.last_render_ms())
    wall = time.perf_counter() - t0
    samples = sum(s.width * s.height * s.spp for s in specs)
    out = {"metric": "samples/s (batched render, 64 instanced-icosphere scenes)",
           "value": samples * args.steps / wall, "unit": "samples/s", "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
           "kernel_ms_avg": float(np.mean(kernel_ms)), "samples_per_step": samples,
           "config": {"workload": "batch_64_instanced_icospheres", "images": len(specs),
                      "triangles": [1e3, 2e6], "sides": [128, 1024], "spp": [min(spp), max(spp)],
                      "depth": specs[0].depth, "launch_tuning": tuning, "autotuner_ms": tune_ms,
                      "bvh_build_s": build_s}}
    if args.sequential:
        total = 0.0
        for a, s, p in zip(arrays, specs, params):
            ds = DeviceScene(dev, a)
            t = Target(ds, s.width, s.height, s.seed)
            t.autotune(p, s.spp, cache_path=TUNING_CACHE)
            t.render(p, s.spp)  # warm-up
            t.set_state(np.zeros((s.width * s.height, 4), np.float32),
                        np.arange(s.width * s.height, dtype=np.uint32) + np.uint32(s.seed * s.width * s.height))
            t.render(p, s.spp)
            total += t.last_render_ms()
            print(f"[sequential] image {len(out.setdefault('sequential_ms', [])) + 1} of {len(specs)}: "
                  f"{t.last_render_ms():.3f} ms", file=sys.stderr, flush=True)
            out["sequential_ms"].append(t.last_render_ms())
            t.destroy()
            ds.destroy()
        out["sequential_kernel_ms"] = total
        out["batch_speedup"] = total / out["kernel_ms_avg"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
