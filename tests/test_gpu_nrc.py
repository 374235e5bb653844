"""GPU: the neural radiance cache's kernels against the CPU oracle
(oracle/nrc_oracle.py) and torch float64 autograd, and what the trained cache
gains over the plain tracer.

* encoding + network inputs: bit for bit (same f32 operation order);
* forward: within BF16 rounding (MFMA accumulates in f32 in its own order,
  which can move an activation by one bf16 step);
* gradients: the kernels' weight and grid gradients against the oracle's
  analytic backward with the same rounding points, and against autograd;
* Adam: two steps against the oracle's update (dense for the MLP, skipping
  zero gradients for the grid);
* training lowers the loss; with termination past the depth limit the
  cache's tracer reproduces the plain render bit for bit; a trained cache
  lowers the RMSE of a 1-spp image against a high-spp reference; frames
  render and train like separate calls."""
import os
import sys

import numpy as np
import pytest

import zig_ornament_amd as zo
from helpers import same_bits, zo_params
from zig_ornament_amd import workloads as W
from zig_ornament_amd.nrc import RECORD_DTYPE, Nrc

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import nrc_oracle as R  # noqa: E402

pytestmark = pytest.mark.gpu


def scene(device, tris=3000, side=48, spp=1, seed=31):
    spec = W.instanced_scene(tris, side, spp, seed=seed)
    return spec, zo.DeviceScene(device, zo.Bvh(W.build_scene(spec)))


def records(info, n, seed):
    rng = np.random.default_rng(seed)
    lo = np.asarray(info["bounds_lo"], np.float64)
    ext = 1 / np.asarray(info["inv_extent"], np.float64)
    r = np.zeros(n, RECORD_DTYPE)
    r["pos"] = lo + rng.uniform(-0.05, 1.05, (n, 3)) * ext  # a few outside the box (clamped)
    d = rng.normal(size=(n, 3))
    r["dir"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    nn = rng.normal(size=(n, 3))
    r["normal"] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
    r["albedo"] = rng.random((n, 3))
    r["roughness"] = rng.random(n) * 0.5
    r["type"] = rng.integers(0, 3, n)
    r["pixel"] = np.arange(n)
    r["weight"] = rng.random((n, 3)) * 1.5
    return r


def random_params(nrc, seed, grid_scale=0.5):
    rng = np.random.default_rng(seed)
    mlp = (rng.normal(size=R.MLP_PARAMS) * 0.15).astype(np.float32)
    gp = nrc.info()["grid_params"]
    grid = (rng.normal(size=gp) * grid_scale).astype(np.float32)
    nrc.set_params(mlp, grid)
    return mlp, grid


@pytest.fixture(scope="module")
def cache(device):
    spec, ds = scene(device)
    nrc = Nrc(ds, log2_table=14, train_batch=1024, train_paths=4096)
    mlp, grid = random_params(nrc, 7)
    info = nrc.info()
    return dict(spec=spec, ds=ds, nrc=nrc, mlp=mlp, grid=grid, info=info,
                table=R.grid_table(info, grid))


def test_params_round_trip(cache):
    mlp, grid = cache["nrc"].get_params()
    assert np.array_equal(mlp, cache["mlp"]) and np.array_equal(grid, cache["grid"])


def test_encoding_and_inputs_are_bit_exact(cache):
    rec = records(cache["info"], 3000, 1)
    print(f"[nrc] levels dense {cache['info']['dense']}", flush=True)
    _, inputs = cache["nrc"].query(rec, with_inputs=True)
    ref = R.inputs(cache["info"], cache["table"], rec)
    assert np.array_equal(inputs, ref), np.argwhere(inputs != ref)[:5]


def test_forward_matches_oracle(cache):
    rec = records(cache["info"], 4000, 2)
    rgb = cache["nrc"].query(rec)
    _, y = R.forward(cache["mlp"], R.inputs(cache["info"], cache["table"], rec))
    err = np.abs(rgb - y[:, :3])
    scale = np.abs(y[:, :3]).max()
    print(f"[nrc] forward max |err| {err.max():.3e} of scale {scale:.3e}", flush=True)
    assert err.max() <= 1e-2 * scale
    assert err.mean() <= 1e-4 * scale


def test_gradients_match_oracle_and_autograd(cache):
    rec = records(cache["info"], 1000, 3)
    loss, mg, gg = cache["nrc"].gradients(rec)
    x = R.inputs(cache["info"], cache["table"], rec)
    xs, y = R.forward(cache["mlp"], x)
    lo, gy = R.loss_and_grad(y, rec["weight"].astype(np.float64))
    g, dx0 = R.backward(cache["mlp"], xs, gy)
    print(f"[nrc] loss gpu {loss:.6f} oracle {lo:.6f}", flush=True)
    assert abs(loss - lo) <= 1e-3 * lo
    assert np.linalg.norm(mg - g) <= 1e-2 * np.linalg.norm(g)
    for l in range(6):
        sl = slice(l * 4096, (l + 1) * 4096) if l < 5 else slice(5 * 4096, None)
        a, b = mg[sl], g[sl]
        assert a @ b / (np.linalg.norm(a) * np.linalg.norm(b)) > 0.999, l
    _, tg, _ = R.torch_reference(cache["mlp"], x, rec["weight"])
    assert np.linalg.norm(mg - tg) <= 3e-2 * np.linalg.norm(tg)
    ggo = R.grid_gradient(cache["info"], rec["pos"], dx0[:, :32])
    assert np.linalg.norm(gg - ggo) <= 1e-2 * np.linalg.norm(ggo)
    touched = np.nonzero(ggo)[0]
    assert np.all(gg[np.setdiff1d(np.arange(gg.size), touched)] == 0)


def test_grid_gradient_at_full_table_size(device):
    """2^19-entry tables, as the 120-frame workload uses them: five densely
    indexed coarse levels whose batch is cut into up to 16 chunks (summed by
    nrc_grid_finish_kernel) and eleven hashed levels of 64 entry ranges each."""
    spec, ds = scene(device, seed=36)
    nrc = Nrc(ds, log2_table=19, train_batch=4096, train_paths=4096)
    mlp, grid = random_params(nrc, 11, grid_scale=0.05)
    info = nrc.info()
    assert info["dense"][:5] == [1] * 5 and not any(info["dense"][5:])
    rec = records(info, 4096, 6)
    loss, mg, gg = nrc.gradients(rec)
    x = R.inputs(info, R.grid_table(info, grid), rec)
    xs, y = R.forward(mlp, x)
    _, gy = R.loss_and_grad(y, rec["weight"].astype(np.float64))
    g, dx0 = R.backward(mlp, xs, gy)
    assert np.linalg.norm(mg - g) <= 1e-2 * np.linalg.norm(g)
    ggo = R.grid_gradient(info, rec["pos"], dx0[:, :32])
    t = 1 << info["log2_table"]
    for l in range(16):
        sl = slice(l * 2 * t, (l + 1) * 2 * t)
        a, b = gg[sl], ggo[sl]
        assert np.linalg.norm(a - b) <= 2e-2 * np.linalg.norm(b), l
        assert np.all(a[b == 0] == 0), l


def test_adam_steps_match_oracle(device):
    """Two Adam steps on different batches: the MLP takes dense steps; grid
    parameters with an exactly-zero gradient in the second batch keep the
    value (and moments) the first step left, although their moments would
    move them under dense Adam."""
    spec, ds = scene(device, seed=32)
    nrc = Nrc(ds, log2_table=12, train_batch=512, learning_rate=1e-3)
    mlp0, grid0 = random_params(nrc, 9)
    info = nrc.info()
    hp = (1e-3, 0.9, 0.99, 1e-15)
    zm, zg = np.zeros(mlp0.size), np.zeros(grid0.size)
    state = dict(mlp=(zm, zm), grid=(zg, zg))
    params = (mlp0, grid0)
    for t, seed in ((1, 4), (2, 5)):
        rec = records(info, 512, seed)
        _, mg, gg = nrc.gradients(rec)
        nrc.step(rec)
        mlp1, grid1 = nrc.get_params()
        p, m, v = R.adam(params[0].astype(np.float64), mg.astype(np.float64), *state["mlp"], t, *hp)
        assert np.allclose(mlp1, p, rtol=0, atol=2e-6), t
        state["mlp"] = (m, v)
        p, m, v = R.adam(params[1].astype(np.float64), gg.astype(np.float64), *state["grid"], t, *hp,
                         skip_zero=True)
        assert np.allclose(grid1, p, rtol=0, atol=2e-6), t
        skipped = gg == 0
        assert np.array_equal(grid1[skipped], params[1][skipped]), t
        if t == 2:  # entries the first step moved and the second skipped
            assert np.count_nonzero(skipped & (state["grid"][0] != 0)) > 100
        state["grid"] = (m, v)
        params = (mlp1, grid1)
    assert nrc.info()["adam_steps"] == 2


def test_training_lowers_the_loss(device):
    spec, ds = scene(device, seed=33)
    nrc = Nrc(ds, log2_table=14, train_batch=2048)
    info = nrc.info()
    rec = records(info, 2048, 5)
    u = R.unit_position(info, rec["pos"])
    rec["weight"] = (0.5 + 0.4 * np.stack([np.sin(4 * u[:, 0]), np.cos(3 * u[:, 1]),
                                           u[:, 2] * rec["albedo"][:, 0]], 1)).astype(np.float32)
    losses = [nrc.step(rec) for _ in range(300)]
    print(f"[nrc] loss {losses[0]:.4f} -> {losses[-1]:.5f}", flush=True)
    assert losses[-1] < 0.1 * losses[0]


def test_no_termination_is_the_plain_render(device):
    """With the query bounce past the depth limit the cache's tracer runs
    the reference estimator alone: the image equals the plain render."""
    spec = W.cornell(width=40, height=32, spp=3)
    ds = zo.DeviceScene(device, zo.Bvh(W.build_scene(spec)))
    nrc = Nrc(ds, log2_table=10, terminate_bounce=spec.depth, train_paths=0)
    a = zo.Target(ds, spec.width, spec.height)
    b = zo.Target(ds, spec.width, spec.height)
    p = zo_params(spec)
    nrc.render(a, p, spec.spp)
    b.render(p, spec.spp)
    assert same_bits(a.accumulation(), b.accumulation())
    assert np.array_equal(a.rng_state(), b.rng_state())
    assert same_bits(a.framebuffer(), b.framebuffer())


@pytest.mark.parametrize("name", ["cornell", "spheres", "instanced"])
def test_primary_queries_trace_like_the_plain_render(device, name):
    """With every surface an emitter no camera path becomes a query, so the
    cache's primary-hit render (nrc_primary_kernel: 8 x 8 wave packets through
    bvh_hit_packet) must equal the plain render bit for bit: camera rays,
    closest hits, sky, emission (textured too), RNG state and accumulation.
    Ragged image sizes leave lanes without a pixel in the last packets."""
    import dataclasses
    spec = {"cornell": W.cornell(width=45, height=37, spp=3),
            "spheres": W.random_spheres(width=50, height=29, spp=3, grid=range(-4, 4)),
            "instanced": W.instanced_icospheres(width=61, height=33, spp=3, levels=2, instances=12,
                                                texture_size=16, box=10.0)}[name]
    spec = dataclasses.replace(spec, materials=[("light", a if a is not None else (1.0, 1.0, 1.0), None)
                                                for _, a, _ in spec.materials])
    ds = zo.DeviceScene(device, zo.Bvh(W.build_scene(spec)))
    nrc = Nrc(ds, log2_table=10, terminate_bounce=0, train_paths=0)
    a = zo.Target(ds, spec.width, spec.height)
    b = zo.Target(ds, spec.width, spec.height)
    p = zo_params(spec)
    nrc.render(a, p, spec.spp)
    b.render(p, spec.spp)
    assert same_bits(a.accumulation(), b.accumulation())
    assert np.array_equal(a.rng_state(), b.rng_state())
    assert same_bits(a.framebuffer(), b.framebuffer())


def test_render_iterations_accumulate_like_single_calls(device):
    """One orn_nrc_render of 3 iterations equals three 1-iteration renders
    with current_iteration 0, 1 and 2: the samples the inference kernel adds
    (queries) and those the path kernel adds (sky, emitters) accumulate as
    State.nextIteration does, and the framebuffer is post-processed once,
    from the 3-sample sum."""
    spec, ds = scene(device, seed=41)
    nrc = Nrc(ds, log2_table=12, terminate_bounce=0, train_paths=0)
    random_params(nrc, 5)
    a = zo.Target(ds, spec.width, spec.height, 3)
    b = zo.Target(ds, spec.width, spec.height, 3)
    nrc.render(a, zo_params(spec), 3)
    for it in range(3):
        nrc.render(b, zo_params(spec, float(it)), 1)
    acc = a.accumulation()
    assert np.all(acc[:, 3] == 3.0)
    assert same_bits(acc, b.accumulation())
    assert np.array_equal(a.rng_state(), b.rng_state())
    assert same_bits(a.framebuffer(), b.framebuffer())


def rmse(img, ref):
    return float(np.sqrt(np.mean((img[:, :3].astype(np.float64) - ref[:, :3]) ** 2)))


def test_trained_cache_lowers_the_error_of_a_one_spp_image(device):
    """What the network gains: on an instanced-icosphere scene, a 1-spp image
    whose paths end in the trained cache at the primary hit is closer to a
    1,024-spp plain reference than a 1-spp plain image."""
    spec, ds = scene(device, tris=20000, side=96, spp=1, seed=34)
    spec.depth = 8
    p = zo_params(spec)
    ref = zo.Target(ds, spec.width, spec.height, 99)
    ref.render(p, 1024)
    reference = ref.accumulation() / 1024.0
    nrc = Nrc(ds, log2_table=16, terminate_bounce=0, train_paths=16384, train_batch=4096)
    for frame in range(40):
        loss = nrc.train(p, frame, steps=4)
    print(f"[nrc] final training loss {loss:.4f}", flush=True)
    a = zo.Target(ds, spec.width, spec.height, 5)
    nrc.render(a, p, 1)
    b = zo.Target(ds, spec.width, spec.height, 5)
    b.render(p, 1)
    e_nrc, e_plain = rmse(a.accumulation(), reference), rmse(b.accumulation(), reference)
    print(f"[nrc] 1-spp RMSE: cache {e_nrc:.4f}, plain {e_plain:.4f}", flush=True)
    assert e_nrc < 0.85 * e_plain  # measured 0.77 on this scene, 0.73 on the 1080p C3 scene


def test_frames_render_and_train_like_separate_calls(device):
    """orn_nrc_frame traces its training paths on a second stream beside the
    render, then takes its steps: two frames give the images and parameters
    of render, train, render, train. Training records are packed in path
    order, so the only difference left is the order of the LDS float sums of
    the grid gradient (a rounding step in a few grid parameters)."""
    spec, ds = scene(device, seed=35)
    p = zo_params(spec)
    # epsilon 1e-3: a parameter whose gradient is at rounding level moves by
    # lr * g / epsilon, not by +-lr in the direction of its rounding noise
    cfg = dict(log2_table=14, terminate_bounce=0, train_paths=4096, train_batch=4096, epsilon=1e-3)
    a = Nrc(ds, **cfg)
    b = Nrc(ds, **cfg)
    ta = zo.Target(ds, spec.width, spec.height, 3)
    tb = zo.Target(ds, spec.width, spec.height, 3)
    ms, la = a.frame(ta, p, 1, 0, steps=2, want_loss=True)
    b.render(tb, p, 1)
    assert ms > 0 and same_bits(ta.accumulation(), tb.accumulation())  # the initial cache in both
    lb = b.train(p, 0, steps=2)
    assert abs(la - lb) <= 1e-5 * lb
    a.frame(ta, p, 1, 1, steps=2)
    b.render(tb, p, 1)
    ia, ib = ta.accumulation(), tb.accumulation()
    d = np.abs(ia - ib)
    print(f"[nrc] frame-1 image: max |diff| {d.max():.3e}, mean {d.mean():.3e}", flush=True)
    assert d.max() <= 1e-3 * np.abs(ib).max() and d.mean() <= 1e-6 * np.abs(ib).mean()
    b.train(p, 1, steps=2)
    ma, ga = a.get_params()
    mb, gb = b.get_params()
    assert a.info()["adam_steps"] == b.info()["adam_steps"] == 4
    assert np.array_equal(ma, mb)
    rel = np.linalg.norm(ga - gb) / np.linalg.norm(gb)
    print(f"[nrc] grid difference after 4 steps: {rel:.2e} (relative norm)", flush=True)
    assert rel <= 1e-6


def test_cache_rejects_bad_arguments(device):
    """Configurations and calls outside the documented ranges return a status
    and leave the cache usable."""
    spec, ds = scene(device, seed=37)
    other_spec, other = scene(device, seed=38)

    INVALID_VALUE, INVALID_HANDLE, SIZE_MISMATCH = 1, 9006, 9007  # include/ornament_hip.h

    def status(fn):
        with pytest.raises(zo.OrnamentError) as e:
            fn()
        return e.value.status

    for bad in (dict(train_batch=100), dict(train_batch=0), dict(log2_table=7), dict(log2_table=25),
                dict(base_resolution=1), dict(max_resolution=8, base_resolution=16),
                dict(learning_rate=0.0), dict(beta1=1.0), dict(epsilon=-1.0)):
        assert status(lambda: Nrc(ds, **bad)) == INVALID_VALUE, bad
    nrc = Nrc(ds, log2_table=10, train_batch=256, train_paths=256)
    p = zo_params(spec)
    t_other = zo.Target(other, spec.width, spec.height)
    assert status(lambda: nrc.render(t_other, p, 1)) == INVALID_HANDLE  # a target of another scene
    t_small = zo.Target(ds, spec.width // 2, spec.height)
    assert status(lambda: nrc.render(t_small, p, 1)) == SIZE_MISMATCH
    mlp, grid = nrc.get_params()
    assert status(lambda: nrc.set_params(mlp[:-1], None)) == SIZE_MISMATCH
    assert status(lambda: nrc.set_params(None, grid[:-1])) == SIZE_MISMATCH
    rec = records(nrc.info(), 300, 7)  # more records than the cache's training capacity (256)
    assert status(lambda: nrc.step(rec)) == INVALID_VALUE
    assert status(lambda: nrc.query(rec[:0])) == INVALID_VALUE
    # still usable
    t = zo.Target(ds, spec.width, spec.height)
    nrc.render(t, p, 1)
    assert np.isfinite(nrc.step(rec[:256]))
    assert nrc.info()["adam_steps"] == 1


def test_training_records_pack_in_path_order_at_scale(device):
    """300,000 training paths (1,172 path workgroups, more than the scan
    workgroup's 1,024 lanes, so each lane sums a run of counts): two caches
    take the same step to the same MLP weights bit for bit, which they can
    only do if both packed the same records in the same order."""
    spec, ds = scene(device, seed=39)
    p = zo_params(spec)
    cfg = dict(log2_table=12, terminate_bounce=0, train_paths=300000, train_batch=65536)
    a, b = Nrc(ds, **cfg), Nrc(ds, **cfg)
    la = a.train(p, 5, steps=1)
    lb = b.train(p, 5, steps=1)
    assert np.isfinite(la) and la > 0 and abs(la - lb) <= 1e-5 * la
    assert np.array_equal(a.get_params()[0], b.get_params()[0])
