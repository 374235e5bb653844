"""The oracle's estimator (pathtracer.wgsl, material.wgsl) against closed
forms that do not depend on how it is written: furnace, depth exhaustion,
RNG draws per material, the sky gradient, and the iteration protocol of
wgpu_backend/path_tracer.zig:212-265."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib as O
from zig_ornament_amd import workloads as W


def enclosure(kind, albedo=(1.0, 1.0, 1.0), extra=(), depth=8, w=16, h=12, spp=4, param=None):
    """Camera at the origin inside a sphere of radius 100 of the given material."""
    spec = W.SceneSpec("enclosure", w, h, spp, depth,
                       camera=((0.0, 0.0, 0.0), (0.0, 0.0, -1.0), (0, 1, 0), w / h, 60.0, 0.0, 10.0))
    spec.materials.append((kind, albedo if kind != "dielectric" else None, param))
    spec.spheres.append(((0.0, 0.0, 0.0), 100.0, 0))
    for e in extra:
        spec.materials.append(e[2])
        spec.spheres.append((e[0], e[1], len(spec.materials) - 1))
    return spec


def render(spec, iterations=None, current=0.0, acc=None, rng=None):
    b = O.build_bvh(spec)
    n = spec.width * spec.height
    acc = np.zeros((n, 4), np.float32) if acc is None else acc
    rng = O.seed_rng(n) if rng is None else rng
    p = O.constant_params(spec, current)
    fb, cur = O.OracleScene(b).render(p, spec.spp if iterations is None else iterations, acc, rng)
    return acc, rng, fb, cur


def advance(state, k):
    s = C.c_uint32(int(state))
    for _ in range(k):
        O.lib().orc_random_u32(C.byref(s))
    return s.value


def test_furnace_returns_exactly_one():
    """Inside a white emitter, white lambertian and glass return exactly 1 per
    path whatever happens, depth exhaustion included."""
    spec = enclosure("light", extra=[((0, 0, -5), 1.0, ("lambertian", (1.0, 1.0, 1.0), None)),
                                     ((2.5, 0, -5), 1.0, ("dielectric", None, 1.5))], depth=3, spp=8)
    acc, _, fb, cur = render(spec)
    assert cur == 8.0
    assert np.all(acc == 8.0)
    assert np.all(fb == 1.0)


@pytest.mark.parametrize("depth,expect", [(1, 0.5), (2, 0.25)])
def test_depth_exhaustion_returns_throughput(depth, expect):
    """pathtracer.wgsl:99-160: when the bounce loop runs out, the throughput
    so far is the sample (not zero). (From depth 3 on, a grazing scatter whose
    chord is shorter than t_min leaves the sphere, as in the reference.)"""
    spec = enclosure("lambertian", albedo=(0.5, 0.5, 0.5), depth=depth, spp=2)
    acc, _, _, _ = render(spec)
    assert np.all(acc[:, :3] == np.float32(2 * expect))


@pytest.mark.parametrize("kind,draws", [("lambertian", 3), ("metal", 3), ("light", 0)])
def test_rng_draws_per_material(kind, draws):
    """2 jitter + 2 lens draws per sample, then lambertian 3, metal 3 (also
    when absorbed), light 0 (depth 1: one shading event per sample)."""
    spec = enclosure(kind, albedo=(0.7, 0.7, 0.7), depth=1, spp=5, param=0.5 if kind == "metal" else None)
    _, rng, _, _ = render(spec)
    n = spec.width * spec.height
    assert all(rng[i] == advance(i, 5 * (4 + draws)) for i in range(0, n, 7))


def test_sky_gradient_closed_form():
    """A miss returns lerp(white, (0.5, 0.7, 1), 0.5 (dir_y + 1)), 4 draws."""
    w, h = 24, 16
    spec = W.SceneSpec("sky", w, h, 1, 4, camera=((0.0, 0.0, 0.0), (0.0, 0.3, -1.0), (0, 1, 0),
                                                  w / h, 70.0, 0.0, 10.0))
    spec.materials.append(("lambertian", (0.5, 0.5, 0.5), None))
    spec.spheres.append(((0.0, 0.0, 50.0), 0.01, 0))  # behind the camera
    acc, rng, _, _ = render(spec)
    cam = O.constant_params(spec).camera
    for i in range(0, w * h, 5):
        s = C.c_uint32(i)
        x, y = i % w, i // w
        u = (x + O.lib().orc_random_f32(C.byref(s))) / (w - 1)
        v = (y + O.lib().orc_random_f32(C.byref(s))) / (h - 1)
        d = (np.array(cam.lower_left_corner, np.float64) + u * np.array(cam.horizontal)
             + v * np.array(cam.vertical) - np.array(cam.origin))
        t = 0.5 * (d[1] / np.linalg.norm(d) + 1.0)
        sky = (1 - t) + t * np.array([0.5, 0.7, 1.0])
        assert np.allclose(acc[i, :3], sky, atol=2e-6)
        assert rng[i] == advance(i, 4)


def test_progressive_render_equals_single_render():
    spec = W.cornell(width=48, height=40, spp=16)
    a1, r1, f1, c1 = render(spec, iterations=16)
    a2, r2, _, c2 = render(spec, iterations=8)
    a2, r2, f2, c2 = render(spec, iterations=8, current=c2, acc=a2, rng=r2)
    assert c1 == c2 == 16.0
    assert a1.tobytes() == a2.tobytes() and r1.tobytes() == r2.tobytes()
    assert f1.tobytes() == f2.tobytes()


def test_first_iteration_overwrites_and_zero_iterations_render_once():
    spec = W.cornell(width=32, height=32, spp=1)
    n = 32 * 32
    a1, _, _, c1 = render(spec, iterations=1)
    garbage = np.full((n, 4), 123.0, np.float32)
    a2, _, _, c2 = render(spec, iterations=0, acc=garbage)
    assert c1 == c2 == 1.0
    assert a1.tobytes() == a2.tobytes()


def test_post_processing_and_flip():
    spec = W.cornell(width=40, height=30, spp=4)
    spec.gamma = 2.2
    acc, _, fb, cur = render(spec)
    ref = np.clip(acc / cur, 0, 1)
    ref[:, :3] = ref[:, :3] ** np.float32(1 / 2.2)
    assert np.allclose(fb, ref, atol=2e-6)
    spec.flip_y = True
    _, _, fb2, _ = render(spec)
    assert np.array_equal(fb2.reshape(30, 40, 4)[::-1], fb.reshape(30, 40, 4))
