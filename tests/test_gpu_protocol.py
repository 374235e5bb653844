"""GPU: the PathTracer interface end to end (scene -> Bvh -> upload ->
render loop -> readback) follows the reference's iteration protocol
(wgpu_backend/path_tracer.zig:198-265, state.zig), checked against the
oracle bit for bit."""
import numpy as np
import pytest

import oracle_lib as O
import zig_ornament_amd as zo
from helpers import same_bits
from zig_ornament_amd import workloads as W

pytestmark = pytest.mark.gpu


def tracer(spec):
    pt = zo.PathTracer(W.build_scene(spec), device=0)
    pt.set_resolution(spec.width, spec.height)
    pt.depth = spec.depth
    return pt


def oracle_run(spec, params, iterations, acc, rng):
    arrays = O.build_bvh(spec)
    p = O.ConstantParams.from_buffer_copy(bytes(params))
    return O.OracleScene(arrays).render(p, iterations, acc, rng)


def test_progressive_renders_accumulate_like_the_reference():
    spec = W.cornell(width=64, height=48, spp=4)
    pt = tracer(spec)
    pt.iterations = 4
    p0 = pt.constant_params()
    pt.render()
    assert pt.current_iteration == 4.0
    pt.render()
    assert pt.current_iteration == 8.0
    n = spec.width * spec.height
    acc, rng = np.zeros((n, 4), np.float32), O.seed_rng(n)
    p0.current_iteration = 0.0
    fb_o, cur = oracle_run(spec, p0, 8, acc, rng)
    assert same_bits(pt.get_accumulation(), acc)
    assert same_bits(pt.get_frame_buffer(), fb_o)
    pt.deinit()


def test_camera_move_resets_accumulation():
    spec = W.cornell(width=40, height=40, spp=2)
    pt = tracer(spec)
    pt.iterations = 2
    pt.render()
    pt.scene.set_look_at((1.0, 1.0, -2.0), (1.0, 1.0, 0.0), (0, 1, 0))  # dirty camera
    pt.render()
    assert pt.current_iteration == 2.0  # State.reset, then two updates
    p = pt.constant_params()
    n = 40 * 40
    # RNG state continues from the first render; accumulation starts over
    acc, rng = np.zeros((n, 4), np.float32), O.seed_rng(n)
    p_first = O.ConstantParams.from_buffer_copy(bytes(p))
    spec_first = W.cornell(width=40, height=40, spp=2)
    arrays = O.build_bvh(spec_first)
    pf = O.constant_params(spec_first)
    O.OracleScene(arrays).render(pf, 2, acc, rng)
    p_first.current_iteration = 0.0
    garbage = np.full((n, 4), 7.0, np.float32)
    O.OracleScene(arrays).render(p_first, 2, garbage, rng)
    assert same_bits(pt.get_accumulation(), garbage)
    pt.deinit()


def test_set_resolution_keeps_state_and_reseeds():
    """path_tracer.zig:198-210 drops the target (new RNG seeds, zeroed
    accumulation) but keeps State.current_iteration."""
    spec = W.cornell(width=32, height=32, spp=2)
    pt = tracer(spec)
    pt.iterations = 2
    pt.render()
    pt.set_resolution(24, 16)
    pt.scene.set_aspect_ratio(24 / 16)  # the app does this too; it dirties the camera
    pt.render()
    assert pt.resolution == (24, 16) and pt.current_iteration == 2.0
    pt.set_resolution(24, 16)  # unchanged size keeps the target
    pt.render()
    assert pt.current_iteration == 4.0
    assert pt.get_frame_buffer().shape == (24 * 16, 4)
    pt.deinit()


def test_zero_iterations_render_once_and_gamma_flip():
    spec = W.cornell(width=32, height=24, spp=1)
    pt = tracer(spec)
    pt.iterations = 0
    pt.gamma = 2.2
    pt.flip_y = True
    p = pt.constant_params()
    pt.render()
    assert pt.current_iteration == 1.0
    n = 32 * 24
    acc, rng = np.zeros((n, 4), np.float32), O.seed_rng(n)
    fb_o, _ = oracle_run(spec, p, 0, acc, rng)
    assert same_bits(pt.get_frame_buffer(), fb_o)
    pt.deinit()


def test_checkpoint_resume(device):
    spec = W.random_spheres(width=96, height=64, spp=8)
    prod = zo.Bvh(W.build_scene(spec))
    ds = zo.DeviceScene(device, prod)
    from helpers import zo_params
    full = zo.Target(ds, spec.width, spec.height)
    full.render(zo_params(spec), 8)
    a = zo.Target(ds, spec.width, spec.height)
    a.render(zo_params(spec), 4)
    acc, rng = a.accumulation(), a.rng_state()
    b = zo.Target(ds, spec.width, spec.height)
    b.set_state(acc, rng)
    b.render(zo_params(spec, current_iteration=4.0), 4)
    assert same_bits(b.accumulation(), full.accumulation())
    assert np.array_equal(b.rng_state(), full.rng_state())


def test_device_errors(device):
    spec = W.cornell(width=16, height=16, spp=1)
    prod = zo.Bvh(W.build_scene(spec))
    ds = zo.DeviceScene(device, prod)
    t = zo.Target(ds, 16, 16)
    from helpers import zo_params
    p = zo_params(W.cornell(width=32, height=16, spp=1))
    with pytest.raises(zo.OrnamentError) as e:
        t.render(p, 1)
    assert e.value.name == "SizeMismatch"
    with pytest.raises(zo.OrnamentError):
        zo.Target(ds, 0, 16)
    assert ds.trace(np.zeros((0, 6), np.float32)).shape == (0,)


def test_tuning_validation_and_defaults(device):
    """orn_target_set_tuning rejects out-of-range fields (node_steps above
    16, workgroup_pool or cost_order above 1) and stores node_steps 0 as its
    default."""
    import ctypes as C
    from zig_ornament_amd import _native as N
    spec = W.cornell(width=16, height=16, spp=1)
    ds = zo.DeviceScene(device, zo.Bvh(W.build_scene(spec)))
    t = zo.Target(ds, 16, 16)
    assert t.tuning()["node_steps"] == 4
    t.set_tuning(node_steps=0)
    assert t.tuning()["node_steps"] == 4
    t.set_tuning(node_steps=16, regroup_lanes=8)
    assert t.tuning()["node_steps"] == 16 and t.tuning()["regroup_lanes"] == 8
    for bad in (dict(node_steps=17), dict(regroup_lanes=0), dict(leaf_regroup_lanes=65),
                dict(bvh_width=3), dict(split_percent=101)):
        with pytest.raises(zo.OrnamentError) as e:
            t.set_tuning(**bad)
        assert e.value.status == 1, bad  # ORN_ERROR_INVALID_VALUE
    raw = N.Tuning()
    assert N.lib().orn_target_get_tuning(t._h, C.byref(raw)) == 0
    raw.workgroup_pool = 2
    assert N.lib().orn_target_set_tuning(t._h, C.byref(raw)) == 1  # ORN_ERROR_INVALID_VALUE
    raw.workgroup_pool, raw.cost_order = 0, 2
    assert N.lib().orn_target_set_tuning(t._h, C.byref(raw)) == 1
    assert t.tuning()["node_steps"] == 16 and t.tuning()["cost_order"] == 0


def test_fractional_iteration_counter_matches_oracle(device):
    """A current_iteration that is not a whole number (never produced by
    State, but accepted by the ABI) keeps the reference's per-sample f32
    increments: such renders run one job per pixel."""
    from helpers import zo_params
    spec = W.cornell(width=24, height=24, spp=5)
    prod = zo.Bvh(W.build_scene(spec))
    ds = zo.DeviceScene(device, prod)
    t = zo.Target(ds, spec.width, spec.height)
    t.set_tuning(samples_per_job=2, split_percent=100)
    for ci in (0.5, 16777214.0):
        p = zo_params(spec, current_iteration=ci)
        t.set_state(np.zeros((24 * 24, 4), np.float32), np.arange(24 * 24, dtype=np.uint32))
        t.render(p, spec.spp)
        acc, rng = np.zeros((24 * 24, 4), np.float32), O.seed_rng(24 * 24)
        fb_o, _ = O.OracleScene(prod).render(O.ConstantParams.from_buffer_copy(bytes(p)),
                                             spec.spp, acc, rng)
        assert same_bits(t.accumulation(), acc), ci
        assert same_bits(t.framebuffer(), fb_o), ci


def test_job_handoff_survives_launch_tag_wrap(device):
    """Pixels handed between jobs are tagged with a 16-bit launch counter
    that wraps after 65,535 renders; 65,600 two-sample renders in two jobs
    each equal one render of all their samples in one job per pixel."""
    from helpers import zo_params
    spec = W.cornell(width=8, height=8, spp=2)
    prod = zo.Bvh(W.build_scene(spec))
    ds = zo.DeviceScene(device, prod)
    a = zo.Target(ds, 8, 8)
    a.set_tuning(samples_per_job=1, split_percent=100)
    renders = 65600
    print(f"[wrap] {renders} renders", flush=True)
    for k in range(renders):
        a.render(zo_params(spec, current_iteration=float(2 * k)), 2, wait=(k % 4096 == 4095))
    a.wait()
    b = zo.Target(ds, 8, 8)
    b.set_tuning(samples_per_job=0)
    b.render(zo_params(spec), 2 * renders)
    assert same_bits(a.accumulation(), b.accumulation())
    assert np.array_equal(a.rng_state(), b.rng_state())
    assert same_bits(a.framebuffer(), b.framebuffer())
