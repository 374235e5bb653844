"""Shared helpers of the test suite: scene construction on both sides, ray
sets, image metrics."""
import numpy as np

import oracle_lib as O
import zig_ornament_amd as zo
from zig_ornament_amd import workloads as W


def both_arrays(spec):
    """Product Bvh arrays and oracle Bvh arrays of one spec."""
    prod = zo.Bvh(W.build_scene(spec))
    orc = O.build_bvh(spec)
    return prod, orc


def ray_set(spec, arrays, n, seed):
    """Seeded world-space rays: half camera rays through random film points,
    half from random points inside the scene bounds in random directions."""
    rng = np.random.default_rng(seed)
    cam = O.constant_params(spec).camera
    o = np.array(cam.origin, np.float64)
    llc = np.array(cam.lower_left_corner, np.float64)
    h = np.array(cam.horizontal, np.float64)
    v = np.array(cam.vertical, np.float64)
    n1 = n // 2
    s = rng.random((n1, 1))
    t = rng.random((n1, 1))
    d1 = llc + s * h + t * v - o
    o1 = np.repeat(o[None, :], n1, 0)
    root = arrays.tlas_nodes[-1]
    if root["node_type"] == 0:
        lo = np.minimum(root["lmin_or_v0"], root["rmin_or_v2"]).astype(np.float64)
        hi = np.maximum(root["lmax_or_v1"], root["rmax_or_v3"]).astype(np.float64)
    else:
        lo = root["lmin_or_v0"].astype(np.float64)
        hi = root["lmax_or_v1"].astype(np.float64)
    # keep huge ground spheres from swallowing the sampling box
    lo = np.maximum(lo, -60.0)
    hi = np.minimum(hi, 60.0)
    n2 = n - n1
    o2 = lo + rng.random((n2, 3)) * (hi - lo)
    d2 = rng.normal(size=(n2, 3))
    d2 *= rng.uniform(0.2, 5.0, size=(n2, 1)) / np.linalg.norm(d2, axis=1, keepdims=True)
    rays = np.concatenate([np.concatenate([o1, d1], 1), np.concatenate([o2, d2], 1)])
    return rays.astype(np.float32)


def rel_mse(x, ref):
    """mean((x - r)^2 / (r^2 + 1e-2)) over RGB."""
    x = np.asarray(x, np.float64)[..., :3]
    r = np.asarray(ref, np.float64)[..., :3]
    return float(np.mean((x - r) ** 2 / (r * r + 1e-2)))


def device_render(device, arrays, spec, iterations=None, seed=0, params=None):
    ds = zo.DeviceScene(device, arrays)
    tgt = zo.Target(ds, spec.width, spec.height, seed)
    p = params if params is not None else zo_params(spec)
    tgt.render(p, spec.spp if iterations is None else iterations)
    return ds, tgt


def zo_params(spec, current_iteration=0.0):
    """ConstantParams for the device ABI, camera from the product's Scene."""
    p = zo._native.ConstantParams()
    p.camera = W.build_scene_camera(spec)
    p.depth, p.width, p.height = spec.depth, spec.width, spec.height
    p.flip_y = 1 if spec.flip_y else 0
    p.inverted_gamma = 1.0 / spec.gamma
    p.ray_cast_epsilon = spec.ray_cast_epsilon
    p.textures_count = len(spec.textures)
    p.current_iteration = current_iteration
    return p


def to_oracle_params(p):
    return O.ConstantParams.from_buffer_copy(bytes(p))


def canon(a):
    """float32 array with every NaN written as the same quiet NaN (x86 and
    gfx950 produce different NaN payloads for the same invalid operation)."""
    a = np.array(a, dtype=np.float32, copy=True)
    a[np.isnan(a)] = np.float32("nan")
    return a


def same_bits(a, b):
    """Bit-for-bit equality of float32 arrays, NaN payloads aside."""
    a, b = canon(a), canon(b)
    return a.shape == b.shape and a.tobytes() == b.tobytes()
