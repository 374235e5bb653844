"""The oracle's RNG and FP32 transcendental functions, pinned against
independent restatements (golden JSON made by tests/golden/make_golden.py),
closed forms and libm in double precision."""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest

import oracle_lib as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_rng_streams_match_golden():
    g = json.load(open(os.path.join(GOLDEN, "rng_streams.json")))
    for state, stream in g["streams"].items():
        u = np.zeros(16, np.uint32)
        f = np.zeros(16, np.float32)
        O.lib().orc_random_stream(int(state), 16, O.ptr(u), O.ptr(f))
        # orc_random_stream returns the draw made from each successive state
        s = C.c_uint32(int(state))
        got = []
        for _ in range(16):
            x = O.lib().orc_random_u32(C.byref(s))
            got.append(x)
        assert got == [p[0] for p in stream], state
        s = C.c_uint32(int(state))
        fb = [np.float32(O.lib().orc_random_f32(C.byref(s))).view(np.uint32) for _ in range(16)]
        assert fb == [p[1] for p in stream], state


def test_rng_can_draw_exactly_one():
    """random.wgsl:21: f32(u32) rounds to nearest, so 1.0 is reachable."""
    g = json.load(open(os.path.join(GOLDEN, "rng_streams.json")))
    s = C.c_uint32(g["state_drawing_one"])
    assert O.lib().orc_random_f32(C.byref(s)) == 1.0


def test_rng_is_an_addition_not_a_multiply():
    """old_state = s + 747796405 + 2891336453 (wraps to s + 3639132858)."""
    for s0 in [0, 7, 0xFFFFFFFF]:
        s = C.c_uint32(s0)
        x = (s0 + 3639132858) & 0xFFFFFFFF
        w = (((x >> ((x >> 28) + 4)) ^ x) * 277803737) & 0xFFFFFFFF
        assert O.lib().orc_random_u32(C.byref(s)) == ((w >> 22) ^ w)


def test_seeding_is_pixel_index():
    """buffers.zig:25-30: state = index (seed 0); seed s adds s * pixel_count."""
    assert np.array_equal(O.seed_rng(100), np.arange(100, dtype=np.uint32))
    assert np.array_equal(O.seed_rng(100, 3), np.arange(300, 400, dtype=np.uint32))


def _eval(op, a, b=None):
    a = np.ascontiguousarray(a, np.float32)
    b = np.zeros_like(a) if b is None else np.ascontiguousarray(b, np.float32)
    out = np.zeros_like(a)
    O.lib().orc_eval_math(op, O.ptr(a), O.ptr(b), O.ptr(out), a.size)
    return out


def _ulp_err(got, ref):
    ref = np.asarray(ref, np.float64)
    ulp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - ref) / ulp


def test_sin_cos_accuracy():
    x = np.linspace(0, 2 * np.pi, 200001, dtype=np.float32)
    es = _ulp_err(_eval(0, x), np.sin(x.astype(np.float64)))
    ec = _ulp_err(_eval(1, x), np.cos(x.astype(np.float64)))
    # absolute error bound near zeros of sin/cos (WGSL: 2^-11 absolute)
    abs_s = np.abs(_eval(0, x) - np.sin(x.astype(np.float64)))
    abs_c = np.abs(_eval(1, x) - np.cos(x.astype(np.float64)))
    assert abs_s.max() < 2e-7 and abs_c.max() < 2e-7
    big = np.abs(np.sin(x.astype(np.float64))) > 0.1
    assert es[big].max() <= 4
    big = np.abs(np.cos(x.astype(np.float64))) > 0.1
    assert ec[big].max() <= 4
    # exact special values
    assert _eval(0, [0.0])[0] == 0.0 and _eval(1, [0.0])[0] == 1.0


def test_pow_accuracy_and_specials():
    rng = np.random.default_rng(0)
    x = rng.random(200000).astype(np.float32)
    for y in [0.33333, 5.0, 1.0, 1 / 2.2]:
        yy = np.full_like(x, y)
        got = _eval(2, x, yy)
        ref = np.power(x.astype(np.float64), np.float64(np.float32(y)))
        err = _ulp_err(got, ref)
        # exp2(y*log2 x): WGSL allows 3 + 2|y log2 x| ulp; this stays far inside
        bound = 3 + 2 * np.abs(np.float64(np.float32(y)) * np.log2(x.astype(np.float64)))
        assert np.all(err <= bound), y
        assert np.median(err) <= 2.0
    assert _eval(2, [0.0], [0.33333])[0] == 0.0
    assert _eval(2, [1.0], [5.0])[0] == 1.0
    assert _eval(2, [0.0], [0.0])[0] == 1.0
    tiny = np.array([1e-40, 1e-30], np.float32)
    assert np.all(np.isfinite(_eval(2, tiny, np.full(2, 0.5, np.float32))))


def test_acos_atan2_accuracy():
    x = np.linspace(-1, 1, 200001, dtype=np.float32)
    e = np.abs(_eval(3, x) - np.arccos(x.astype(np.float64)))
    assert e.max() < 5e-7
    assert _eval(3, [1.0])[0] == 0.0
    assert abs(_eval(3, [-1.0])[0] - math.pi) < 1e-6
    rng = np.random.default_rng(1)
    y = rng.uniform(-1, 1, 200000).astype(np.float32)
    xx = rng.uniform(-1, 1, 200000).astype(np.float32)
    e = np.abs(_eval(4, y, xx) - np.arctan2(y.astype(np.float64), xx.astype(np.float64)))
    assert e.max() < 8e-7
    # quadrant axes
    got = _eval(4, [0, 1, 0, -1], [1, 0, -1, 0])
    assert np.allclose(got, [0, math.pi / 2, math.pi, -math.pi / 2], atol=3e-7)


def test_unit_sphere_and_disk_draws():
    """random.wgsl:36-87: |p| <= 1, r^3 uniform, theta uniform in [0, pi]
    (the reference's non-uniform sphere), disk r^2 uniform."""
    n = 40000
    pts = np.zeros((n, 3), np.float32)
    disk = np.zeros((n, 3), np.float32)
    s = C.c_uint32(99)
    buf = (C.c_float * 3)()
    for i in range(n):
        O.lib().orc_random_in_unit_sphere(C.byref(s), buf)
        pts[i] = list(buf)
    for i in range(n):
        O.lib().orc_random_in_unit_disk(C.byref(s), buf)
        disk[i] = list(buf)
    r = np.linalg.norm(pts.astype(np.float64), axis=1)
    assert r.max() <= 1.0 + 1e-6
    from scipy import stats
    assert stats.kstest(r ** 3, "uniform").pvalue > 1e-3
    theta = np.arccos(np.clip(pts[:, 2] / np.maximum(r, 1e-30), -1, 1))
    assert stats.kstest(theta / np.pi, "uniform").pvalue > 1e-3
    rd = np.linalg.norm(disk[:, :2].astype(np.float64), axis=1)
    assert np.all(disk[:, 2] == 0)
    assert stats.kstest(rd ** 2, "uniform").pvalue > 1e-3
