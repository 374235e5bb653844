"""GPU: batched rendering (many scenes in one persistent launch replayed
from a HIP graph) gives every image exactly what a separate target gives,
and the autotuner leaves a target's state untouched.

A batch is n independent PathTracer.render() calls
(wgpu_backend/path_tracer.zig:252-265) executed together; since each
pixel's RNG chain and accumulation are its own, batching must not change a
single bit of any image."""
import json

import numpy as np
import pytest

import oracle_lib as O
import zig_ornament_amd as zo
from helpers import same_bits, zo_params
from zig_ornament_amd import workloads as W

pytestmark = pytest.mark.gpu


def mixed_specs():
    """Scenes of three feature classes (Cornell: meshes and spheres; random
    spheres: spheres only; instanced icospheres: textured meshes), odd sizes
    with partial edge tiles, flip_y, gamma and RNG seeds."""
    a = W.cornell(width=70, height=45, spp=3)
    b = W.random_spheres(width=96, height=54, spp=2, depth=6)
    b.flip_y, b.gamma, b.seed = True, 2.2, 3
    c = W.instanced_scene(3000, 48, 2, seed=11)
    d = W.instanced_scene(60000, 33, 1, seed=12)
    d.seed = 1
    e = W.cornell(width=17, height=130, spp=1)
    return [a, b, c, d, e]


def batch_of(device, specs):
    arrays = [zo.Bvh(W.build_scene(s)) for s in specs]
    batch = zo.Batch(device, [(a, s.width, s.height, s.seed) for a, s in zip(arrays, specs)])
    return arrays, batch


def test_batch_matches_separate_targets_over_graph_replays(device):
    specs = mixed_specs()
    arrays, batch = batch_of(device, specs)
    params = [zo_params(s) for s in specs]
    iters = [s.spp for s in specs]
    print(f"[batch] {len(specs)} images, first render", flush=True)
    batch.render(params, iters)
    # a second render through the same graph continues every image
    params2 = [zo_params(s, current_iteration=float(max(s.spp, 1))) for s in specs]
    iters2 = [2, 1, 3, 1, 2]
    batch.render(params2, iters2)
    for i, s in enumerate(specs):
        ds = zo.DeviceScene(device, arrays[i])
        t = zo.Target(ds, s.width, s.height, s.seed)
        t.render(params[i], iters[i])
        t.render(params2[i], iters2[i])
        assert same_bits(batch.accumulation(i), t.accumulation()), s.name
        assert np.array_equal(batch.rng_state(i), t.rng_state()), s.name
        assert same_bits(batch.framebuffer(i), t.framebuffer()), s.name
    batch.destroy()


@pytest.mark.parametrize("tuning", [dict(cost_order=1),
                                    dict(cost_order=1, samples_per_job=2, split_percent=50, regroup_lanes=16)])
def test_cost_ordered_batch_matches_separate_targets(device, tuning):
    """With cost_order 1 the batch graph orders each band's queue by the
    pixel costs its previous replay recorded (the first replay has none and
    keeps the work-list order). Three replays in a row must each leave every
    image as a separate target leaves it."""
    specs = mixed_specs()
    arrays, batch = batch_of(device, specs)
    batch.set_tuning(**tuning)
    targets = []
    for i, s in enumerate(specs):
        ds = zo.DeviceScene(device, arrays[i])
        targets.append((ds, zo.Target(ds, s.width, s.height, s.seed)))
    for k in range(3):
        print(f"[batch] cost-ordered replay {k}, {tuning}", flush=True)
        params = [zo_params(s, current_iteration=float(k * max(s.spp, 1))) for s in specs]
        iters = [s.spp for s in specs]
        batch.render(params, iters)
        for i, s in enumerate(specs):
            t = targets[i][1]
            t.render(params[i], iters[i])
            assert same_bits(batch.accumulation(i), t.accumulation()), (k, s.name)
            assert np.array_equal(batch.rng_state(i), t.rng_state()), (k, s.name)
            assert same_bits(batch.framebuffer(i), t.framebuffer()), (k, s.name)
    batch.destroy()


def test_batch_matches_oracle_image(device, oracle):
    specs = [W.instanced_scene(2000, 24, 2, seed=21), W.cornell(width=20, height=20, spp=2)]
    arrays, batch = batch_of(device, specs)
    params = [zo_params(s) for s in specs]
    batch.render(params, [s.spp for s in specs])
    for i, s in enumerate(specs):
        n = s.width * s.height
        acc, rng = np.zeros((n, 4), np.float32), O.seed_rng(n, s.seed)
        fb, _ = O.OracleScene(O.build_bvh(s)).render(
            O.ConstantParams.from_buffer_copy(bytes(params[i])), s.spp, acc, rng)
        assert same_bits(batch.framebuffer(i), fb), s.name
        assert np.array_equal(batch.rng_state(i), rng), s.name


@pytest.mark.parametrize("tuning", [
    dict(regroup_lanes=8, leaf_regroup_lanes=1, samples_per_job=1, split_percent=100),
    dict(regroup_lanes=64, leaf_regroup_lanes=32, workgroups_per_cu=1, samples_per_job=0),
    dict(regroup_lanes=16, leaf_regroup_lanes=4, node_steps=16),
    dict(regroup_lanes=48, node_steps=1, samples_per_job=2, split_percent=50),
])
def test_batch_tunings_give_the_same_images(device, tuning):
    specs = mixed_specs()[:4]
    arrays, batch = batch_of(device, specs)
    params = [zo_params(s) for s in specs]
    iters = [s.spp for s in specs]
    batch.render(params, iters)
    ref = [batch.accumulation(i) for i in range(len(specs))]
    for i in range(len(specs)):  # back to the seeded state
        n = specs[i].width * specs[i].height
        batch.set_state(i, np.zeros((n, 4), np.float32),
                        np.arange(n, dtype=np.uint32) + np.uint32(specs[i].seed * n))
    batch.set_tuning(**tuning)
    assert batch.tuning()["regroup_lanes"] == tuning["regroup_lanes"]
    assert batch.tuning()["node_steps"] == tuning.get("node_steps", 4)
    batch.render(params, iters)
    for i in range(len(specs)):
        assert same_bits(batch.accumulation(i), ref[i]), specs[i].name


def test_batch_rejects_bad_arguments(device):
    specs = mixed_specs()[:2]
    arrays, batch = batch_of(device, specs)
    p = [zo_params(s) for s in specs]
    p[1].width += 1
    with pytest.raises(zo.OrnamentError) as e:
        batch.render(p, [1, 1])
    assert e.value.name == "SizeMismatch"
    out = np.zeros((specs[0].width * specs[0].height, 4), np.float32)
    assert zo._native.lib().orn_batch_get_framebuffer(batch.handle, 7, out.ctypes.data,
                                                      len(out)) == 1  # item out of range
    assert zo._native.lib().orn_batch_get_framebuffer(batch.handle, 0, out.ctypes.data,
                                                      len(out) - 1) == 9007  # too small


def test_target_autotune_keeps_state_and_uses_cache(device, tmp_path):
    spec = W.cornell(width=64, height=64, spp=2)
    arrays = zo.Bvh(W.build_scene(spec))
    ds = zo.DeviceScene(device, arrays)
    t = zo.Target(ds, spec.width, spec.height, 0)
    p = zo_params(spec)
    t.render(p, 2)
    before = (t.accumulation(), t.rng_state(), t.framebuffer())
    cache = str(tmp_path / "tune.json")
    chosen, ms = t.autotune(p, 2, cache_path=cache)
    print(f"[autotune] {chosen} {ms:.3f} ms", flush=True)
    assert ms > 0 and t.tuning() == chosen and chosen["bvh_width"] == 2
    after = (t.accumulation(), t.rng_state(), t.framebuffer())
    assert same_bits(before[0], after[0]) and np.array_equal(before[1], after[1])
    assert same_bits(before[2], after[2])
    entries = json.load(open(cache))
    assert len(entries) == 1
    key = next(iter(entries))
    assert key.startswith(device.info()[0].split("(")[-1].rstrip(")"))
    # a second target of the same class takes the cached entry without timing
    t2 = zo.Target(ds, spec.width, spec.height, 0)
    chosen2, ms2 = t2.autotune(p, 2, cache_path=cache)
    assert chosen2 == chosen and abs(ms2 - entries[key]["ms"]) < 1e-3
    assert entries[key]["node_steps"] == chosen["node_steps"]
    # an entry written before node_steps existed reads as the default (4)
    del entries[key]["node_steps"]
    old = str(tmp_path / "old.json")
    json.dump(entries, open(old, "w"))
    chosen_old, _ = zo.Target(ds, spec.width, spec.height, 0).autotune(p, 2, cache_path=old)
    assert chosen_old == dict(chosen, node_steps=4)
    # and renders afterwards still match the untuned render
    t2.render(p, 2)
    t3 = zo.Target(ds, spec.width, spec.height, 0)
    t3.render(p, 2)
    assert same_bits(t2.accumulation(), t3.accumulation())


def test_batch_autotune_keeps_state(device):
    specs = mixed_specs()[:3]
    arrays, batch = batch_of(device, specs)
    params = [zo_params(s) for s in specs]
    iters = [s.spp for s in specs]
    batch.render(params, iters)
    before = [batch.accumulation(i) for i in range(3)]
    chosen, ms = batch.autotune(params, [1, 1, 1])
    assert ms > 0 and batch.tuning() == chosen
    for i in range(3):
        assert same_bits(batch.accumulation(i), before[i])
