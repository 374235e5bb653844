"""Closed-form checks of the oracle's camera and primitive tests
(camera.wgsl, bvh.wgsl:139-249, transform.wgsl)."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib as O

f3 = lambda *v: (C.c_float * 3)(*[float(x) for x in v])  # noqa: E731


def test_camera_basis_closed_form():
    cam = O.camera((13, 2, 3), (0, 0, 0), (0, 1, 0), 16 / 9, 20.0, 0.1, 10.0)
    w = np.array(cam.w, np.float64)
    u = np.array(cam.u, np.float64)
    v = np.array(cam.v, np.float64)
    look = np.array([13, 2, 3], np.float64)
    assert np.allclose(w, look / np.linalg.norm(look), atol=1e-6)
    assert abs(u @ w) < 1e-6 and abs(v @ w) < 1e-6 and abs(u @ v) < 1e-6
    h = math.tan(math.radians(20.0) / 2)
    assert np.allclose(np.array(cam.vertical), 10.0 * 2 * h * v, rtol=1e-5)
    assert np.allclose(np.array(cam.horizontal), 10.0 * 2 * h * 16 / 9 * u, rtol=1e-5)
    assert cam.lens_radius == pytest.approx(0.05)
    # the film centre is focus_dist along -w: llc + h/2 + v/2 - o = -focus * w
    centre = (np.array(cam.lower_left_corner) + np.array(cam.horizontal) / 2
              + np.array(cam.vertical) / 2 - np.array(cam.origin))
    assert np.allclose(centre, -10.0 * w, atol=1e-4)


def test_camera_ray_draws_lens_even_without_aperture():
    """camera.wgsl:13: two disk draws although lens_radius == 0; the ray
    direction is not normalised."""
    cam = O.camera((0, 0, 0), (0, 0, -1), (0, 1, 0), 1.0, 90.0, 0.0, 10.0)
    s = C.c_uint32(5)
    o, d = (C.c_float * 3)(), (C.c_float * 3)()
    O.lib().orc_camera_get_ray(C.byref(cam), 0.5, 0.5, C.byref(s), o, d)
    s2 = C.c_uint32(5)
    O.lib().orc_random_u32(C.byref(s2))
    O.lib().orc_random_u32(C.byref(s2))
    assert s.value == s2.value
    assert list(o) == [0.0, 0.0, 0.0]
    assert np.allclose(list(d), [0, 0, -10.0], atol=1e-5)


def test_safe_invdir():
    out = (C.c_float * 3)()
    O.lib().orc_safe_invdir(f3(0.0, -0.0, -1e-6), out)
    assert list(out) == [np.float32(1e5), np.float32(1e5), np.float32(-1e5)] or \
        np.allclose(list(out), [1e5, 1e5, -1e5])
    O.lib().orc_safe_invdir(f3(2.0, -4.0, 1e-4), out)
    assert np.allclose(list(out), [0.5, -0.25, 1e4])


def test_aabb_slab():
    inv, ox, out = (C.c_float * 3)(), (C.c_float * 3)(), (C.c_float * 2)()
    o = np.array([-5.0, 0.5, 0.5], np.float32)
    d = np.array([1.0, 0.0, 0.0], np.float32)
    O.lib().orc_safe_invdir(f3(*d), inv)
    for k in range(3):
        ox[k] = np.float32(-o[k]) * inv[k]
    O.lib().orc_aabb_hit(f3(0, 0, 0), f3(1, 1, 1), inv, ox, 0.001, 1e30, out)
    assert out[0] == pytest.approx(5.0) and out[1] == pytest.approx(6.0)
    # t_max clips: a box beyond t_max misses
    O.lib().orc_aabb_hit(f3(0, 0, 0), f3(1, 1, 1), inv, ox, 0.001, 4.0, out)
    assert out[0] > out[1]


def test_triangle_edges_inclusive_and_t_bounds():
    v0, v1, v2 = f3(0, 0, 0), f3(1, 0, 0), f3(0, 1, 0)
    uv = (C.c_float * 2)()
    d = f3(0, 0, -1)
    t = O.lib().orc_triangle_hit(f3(0.25, 0.25, 1), d, v0, v1, v2, 0.001, 1e30, uv)
    assert t == 1.0 and list(uv) == [0.25, 0.25]
    # vertices and edges count as hits (u, v in [0, 1], u + v <= 1)
    for p in [(0, 0), (1, 0), (0, 1), (0.5, 0.5), (0.5, 0)]:
        assert O.lib().orc_triangle_hit(f3(p[0], p[1], 1), d, v0, v1, v2, 0.001, 1e30, uv) == 1.0
    # just outside
    assert O.lib().orc_triangle_hit(f3(0.6, 0.6, 1), d, v0, v1, v2, 0.001, 1e30, uv) == np.float32(1e30)
    # t exactly at t_min is a hit, t > t_max a miss (bvh.wgsl:220)
    assert O.lib().orc_triangle_hit(f3(0.25, 0.25, 0.5), d, v0, v1, v2, 0.5, 1e30, uv) == 0.5
    assert O.lib().orc_triangle_hit(f3(0.25, 0.25, 1), d, v0, v1, v2, 0.001, 0.5, uv) == np.float32(0.5)


def test_sphere_roots():
    # outside: near root
    assert O.lib().orc_sphere_hit(f3(0, 0, 5), f3(0, 0, -1), 0.001, 1e30) == 4.0
    # inside: far root
    assert O.lib().orc_sphere_hit(f3(0, 0, 0), f3(0, 0, -2), 0.001, 1e30) == 0.5
    # miss returns t_max
    assert O.lib().orc_sphere_hit(f3(0, 2, 5), f3(0, 0, -1), 0.001, 7.0) == 7.0
    # unnormalised directions scale t
    assert O.lib().orc_sphere_hit(f3(0, 0, 5), f3(0, 0, -4), 0.001, 1e30) == 1.0


def test_matrix_helpers():
    m = (C.c_float * 16)(*np.eye(4, dtype=np.float32).ravel())
    inv = (C.c_float * 16)()
    assert O.lib().orc_mat_inverse(m, inv) == 0
    assert list(inv) == list(np.eye(4, dtype=np.float32).ravel())
    # S(r) T(c): inverse is exactly S(1/r) T(-c/r) when those are representable
    t, lo, hi = (C.c_float * 16)(), (C.c_float * 4)(), (C.c_float * 4)()
    O.lib().orc_create_sphere(O.fa([1.0, 2.0, -4.0, 1.0]), 2.0, t, lo, hi)
    O.lib().orc_mat_inverse(t, inv)
    ref = np.eye(4)
    ref[:3, :3] *= 0.5
    ref[3, :3] = [-0.5, -1.0, 2.0]
    assert np.array_equal(np.array(list(inv)).reshape(4, 4), ref)
    assert list(lo) == [-1.0, 0.0, -6.0, 1.0] and list(hi) == [3.0, 4.0, -2.0, 1.0]
    # rotationBetweenVectors(n, +y) maps n to +y (row vectors: n * R)
    for n in [(1, 0, 0), (-1, 0, 0), (0, -1, 0), (0, 0, 1), (0.6, 0.8, 0)]:
        r = (C.c_float * 16)()
        O.lib().orc_rotation_between_vectors(O.fa(list(n) + [0.0]), O.fa([0, 1, 0, 0]), r)
        R = np.array(list(r), np.float64).reshape(4, 4)
        assert np.allclose(np.array(n, np.float64) @ R[:3, :3], [0, 1, 0], atol=1e-6), n
        assert np.allclose(R[:3, :3] @ R[:3, :3].T, np.eye(3), atol=1e-6)
