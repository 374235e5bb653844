"""GPU parity: the gfx950 kernels against the CPU oracle on the same inputs.

Bar: bit-exact closest hits (IDs, t, barycentrics) and bit-exact
accumulation buffers at equal spp (the kernels and the oracle define every
floating-point operation identically); relMSE <= 1e-4 is asserted as well.
Closest hits are also checked against the FP64 brute-force oracle: equal IDs
on every ray it does not flag as an edge/tie case, t within 1e-5 relative.
"""
import os
import sys

import numpy as np
import pytest

import oracle_lib as O
import zig_ornament_amd as zo
from helpers import (both_arrays, device_render, ray_set, rel_mse, same_bits, to_oracle_params,
                     zo_params)
from zig_ornament_amd import workloads as W

pytestmark = pytest.mark.gpu


def small_specs():
    return {
        "C1": W.cornell(),
        "C2s": W.random_spheres(width=160, height=90, spp=8),
        "C3s": W.instanced_icospheres(width=96, height=54, spp=4, levels=3, instances=16,
                                      texture_size=64, box=12.0),
    }


@pytest.mark.parametrize("width", [2, 4])
@pytest.mark.parametrize("name", ["C1", "C2s", "C3s"])
def test_trace_rays_match_oracle(device, name, width):
    """Both node widths: the reference's BVH2 and the derived BVH4."""
    spec = small_specs()[name]
    prod, orc = both_arrays(spec)
    rays = ray_set(spec, prod, 20000, seed=7)
    print(f"[trace] {name} width {width}: {len(rays)} rays", flush=True)
    ds = zo.DeviceScene(device, prod)
    ds.set_trace_width(width)
    g = ds.trace(rays, spec.ray_cast_epsilon)
    o = O.OracleScene(orc).trace(rays, spec.ray_cast_epsilon)
    assert np.array_equal(g["hit"], o["hit"])
    h = o["hit"] == 1
    assert h.mean() > 0.1
    for f in ["material_index", "node_type", "inverted_transform_id"]:
        assert np.array_equal(g[f][h], o[f][h]), f
    assert np.array_equal(g["t"][h].view(np.uint32), o["t"][h].view(np.uint32))
    m = h & (o["node_type"] == 2)
    assert np.array_equal(g["tri_id"][m], o["tri_id"][m])
    assert np.array_equal(g["uv"][m].view(np.uint32), o["uv"][m].view(np.uint32))


def tiled_camera_rays(spec, width, height):
    """Pinhole rays through the pixel centres of a width x height film, in
    8 x 8 tiles (64 consecutive rays per tile, as a wave traces them)."""
    cam = O.constant_params(spec).camera
    o = np.array(cam.origin, np.float64)
    llc = np.array(cam.lower_left_corner, np.float64)
    h = np.array(cam.horizontal, np.float64)
    v = np.array(cam.vertical, np.float64)
    out = []
    for ty in range(0, height, 8):
        for tx in range(0, width, 8):
            for k in range(64):
                x, y = tx + k % 8, ty + k // 8
                d = llc + (x + 0.5) / width * h + (y + 0.5) / height * v - o
                out.append(np.concatenate([o, d]))
    return np.asarray(out, np.float32)


@pytest.mark.parametrize("name", ["C1", "C2s", "C3s"])
def test_packet_traversal_matches_per_ray(device, name):
    """orn_trace_rays_packet traces each wave's 64 rays together; every ray
    must still get its own traversal's closest hit bit for bit, for coherent
    camera tiles and for scattered rays (including a partial last wave)."""
    spec = small_specs()[name]
    prod, orc = both_arrays(spec)
    ds = zo.DeviceScene(device, prod)
    for label, rays in (("tiles", tiled_camera_rays(spec, 128, 72)),
                        ("scattered", ray_set(spec, prod, 20001, seed=9))):
        print(f"[trace] {name} packets, {label}: {len(rays)} rays", flush=True)
        g = ds.trace(rays, spec.ray_cast_epsilon, packet=True)
        r = ds.trace(rays, spec.ray_cast_epsilon)
        o = O.OracleScene(orc).trace(rays, spec.ray_cast_epsilon)
        assert np.array_equal(r["hit"], o["hit"])
        assert g.tobytes() == r.tobytes(), label
        assert o["hit"].mean() > 0.1


@pytest.mark.parametrize("name", ["C1", "C2s", "C3s"])
def test_trace_rays_match_fp64_brute_force(device, name):
    """IDs equal on every ray the FP64 oracle does not flag (edge, tangency,
    t_min or tie within 1e-5); t within 1e-5 relative wherever the FP32
    operation sequence can resolve that (its forward error bound), and within
    that bound everywhere."""
    spec = small_specs()[name]
    prod, _ = both_arrays(spec)
    rays = ray_set(spec, prod, 20000, seed=11)
    g = zo.DeviceScene(device, prod).trace(rays, spec.ray_cast_epsilon)
    bf = O.OracleScene(prod).brute_force(rays, spec.ray_cast_epsilon, rel_eps=1e-5)
    amb = bf[:, 4] == 1
    ok = ~amb
    print(f"[bf] {name}: ambiguous {amb.sum()} of {len(rays)}", flush=True)
    assert amb.mean() < 0.03
    assert np.array_equal(g["hit"][ok], bf[ok, 0].astype(np.uint32))
    h = ok & (bf[:, 0] == 1)
    assert np.array_equal(g["inverted_transform_id"][h] // 2, bf[h, 2].astype(np.uint32))
    tri = h & (g["node_type"] == 2)
    assert np.array_equal(g["tri_id"][tri] // 3, bf[tri, 3].astype(np.uint32))
    rel = np.abs(g["t"][h].astype(np.float64) - bf[h, 1]) / np.abs(bf[h, 1])
    bound = bf[h, 6]
    resolvable = bound <= 1e-5
    print(f"[bf] {name}: hits {h.sum()}, within 1e-5: {(rel <= 1e-5).mean():.4f}, "
          f"FP32-resolvable {resolvable.mean():.4f}", flush=True)
    assert np.all(rel[resolvable] <= 1e-5)
    assert np.all(rel <= np.maximum(1e-5, bound))


@pytest.mark.parametrize("tuning", [dict(bvh_width=4), dict(bvh_width=2),
                                    dict(bvh_width=4, regroup_lanes=1, workgroups_per_cu=1),
                                    dict(bvh_width=2, regroup_lanes=64, workgroups_per_cu=3),
                                    dict(samples_per_job=1, split_percent=100, regroup_lanes=8),
                                    dict(samples_per_job=0),
                                    dict(samples_per_job=3, split_percent=60, workgroups_per_cu=1),
                                    dict(regroup_lanes=32, leaf_regroup_lanes=8, samples_per_job=4,
                                         split_percent=100),
                                    dict(node_steps=1), dict(node_steps=16, leaf_regroup_lanes=1),
                                    dict(node_steps=6, regroup_lanes=8, samples_per_job=2, split_percent=100),
                                    dict(workgroup_pool=1),
                                    dict(workgroup_pool=1, regroup_lanes=8, samples_per_job=3, split_percent=100,
                                         node_steps=2, workgroups_per_cu=1)])
@pytest.mark.parametrize("name", ["C1", "C2s", "C3s"])
def test_image_matches_oracle(device, name, tuning):
    """Any launch tuning (node width, regrouping, persistent grid size, job
    size, internal-node steps between round checks, rays pooled across a
    workgroup's waves, which the spheres scene runs and the mesh scenes
    ignore) renders the oracle's image bit for bit."""
    spec = small_specs()[name]
    prod, _ = both_arrays(spec)
    n = spec.width * spec.height
    print(f"[image] {name}: {spec.width}x{spec.height} {spec.spp} spp, {tuning}", flush=True)
    ds = zo.DeviceScene(device, prod)
    tgt = zo.Target(ds, spec.width, spec.height)
    tgt.set_tuning(**tuning)
    tgt.render(zo_params(spec), spec.spp)
    acc_g, fb_g, rng_g = tgt.accumulation(), tgt.framebuffer(), tgt.rng_state()
    p = to_oracle_params(zo_params(spec))
    acc_o = np.zeros((n, 4), np.float32)
    rng_o = O.seed_rng(n)
    fb_o, cur = O.OracleScene(prod).render(p, spec.spp, acc_o, rng_o)
    assert cur == spec.spp
    exact = np.all((acc_g == acc_o) | (np.isnan(acc_g) & np.isnan(acc_o)), axis=1)
    print(f"[image] {name}: exact pixels {exact.mean():.6f}, "
          f"relMSE {rel_mse(np.nan_to_num(acc_g) / cur, np.nan_to_num(acc_o) / cur):.3e}",
          flush=True)
    assert rel_mse(np.nan_to_num(acc_g) / cur, np.nan_to_num(acc_o) / cur) <= 1e-4
    assert exact.mean() == 1.0
    assert np.array_equal(rng_g, rng_o)
    assert same_bits(fb_g, fb_o)


@pytest.mark.parametrize("tuning", [dict(cost_order=1),
                                    dict(cost_order=1, samples_per_job=3, split_percent=100, regroup_lanes=8),
                                    dict(cost_order=1, samples_per_job=2, split_percent=60, node_steps=6),
                                    dict(cost_order=1, bvh_width=4),
                                    dict(cost_order=1, workgroup_pool=1, samples_per_job=3, split_percent=100,
                                         regroup_lanes=8, workgroups_per_cu=1)])
@pytest.mark.parametrize("name", ["C1", "C2s", "C3s", "C1tiny"])
def test_cost_ordered_renders_match_oracle(device, name, tuning):
    """With cost_order 1 the first render records each pixel's cost and
    every later one queues the pixels by the previous render's costs (a
    permutation of each band's queue, built on the GPU). Three progressive
    renders must each equal the oracle bit for bit: a pixel queued twice or
    never would leave a wrong accumulation or RNG state. C1tiny has two
    partial tiles, so six of the eight bands are empty."""
    spec = W.cornell(width=20, height=12, spp=3) if name == "C1tiny" else small_specs()[name]
    prod, _ = both_arrays(spec)
    n = spec.width * spec.height
    ds = zo.DeviceScene(device, prod)
    tgt = zo.Target(ds, spec.width, spec.height)
    tgt.set_tuning(**tuning)
    with pytest.raises(zo.OrnamentError):
        tgt.pixel_costs()  # nothing recorded yet
    acc_o = np.zeros((n, 4), np.float32)
    rng_o = O.seed_rng(n)
    oracle = O.OracleScene(prod)
    for k in range(3):
        print(f"[ordered] {name} render {k}, {tuning}", flush=True)
        params = zo_params(spec, current_iteration=float(k * spec.spp))
        tgt.render(params, spec.spp)
        costs = tgt.pixel_costs()
        assert np.all(costs > 0)
        fb_o, _ = oracle.render(to_oracle_params(params), spec.spp, acc_o, rng_o)
        assert same_bits(tgt.accumulation(), acc_o)
        assert np.array_equal(tgt.rng_state(), rng_o)
        assert same_bits(tgt.framebuffer(), fb_o)


def test_activation_bf16_rounding_is_round_to_nearest_even(device):
    """The radiance cache's MLP rounds its relu'd activations to bf16 with
    the hardware conversion (orn_eval_math op 5). For every non-NaN float it
    must equal the oracle's round to nearest even, (u + 0x7fff + ((u >> 16)
    & 1)) >> 16: random bit patterns of every exponent, exact ties with an
    even and an odd kept half, the values next to ties, subnormals, +-0,
    +-inf and the largest floats (which round up to inf)."""
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
    import nrc_oracle as R
    rng = np.random.default_rng(17)
    bits = rng.integers(0, 1 << 32, 1 << 18, dtype=np.uint64).astype(np.uint32)
    hi = rng.integers(0, 1 << 16, 4096, dtype=np.uint64).astype(np.uint32) << 16
    ties = np.concatenate([hi | 0x8000, hi | 0x7FFF, hi | 0x8001])
    special = np.array([0, 0x80000000, 1, 0x7FFFFF, 0x80000001, 0x807FFFFF, 0x7F800000, 0xFF800000,
                        0x7F7FFFFF, 0xFF7FFFFF, 0x7F7F8000, 0x7F7E8000, 0x00008000, 0x00018000],
                       np.uint32)
    u = np.concatenate([bits, ties, special])
    a = u.view(np.float32)
    a = a[~np.isnan(a)]
    g = device.eval_math(5, a)
    o = R.bf16(a)
    assert np.array_equal(g.view(np.uint32), o.view(np.uint32)), a[g.view(np.uint32) != o.view(np.uint32)][:8]


def test_math_functions_bit_exact(device):
    rng = np.random.default_rng(5)
    n = 1 << 16
    cases = {
        0: (rng.uniform(0, 2 * np.pi, n), None),
        1: (rng.uniform(0, 2 * np.pi, n), None),
        2: (rng.uniform(0, 1, n), rng.choice([0.33333, 5.0, 1.0, 1 / 2.2], n)),
        3: (rng.uniform(-1, 1, n), None),
        4: (rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)),
    }
    for op, (a, b) in cases.items():
        a = a.astype(np.float32)
        b = np.zeros_like(a) if b is None else b.astype(np.float32)
        g = device.eval_math(op, a, b)
        o = np.zeros_like(a)
        O.lib().orc_eval_math(op, O.ptr(a), O.ptr(b), O.ptr(o), n)
        assert np.array_equal(g.view(np.uint32), o.view(np.uint32)), op
