"""GPU at the benchmark sizes (random spheres 1920x1080 64 spp depth 16;
Cornell 256x256 16 spp; instanced icospheres 1920x1080 32 spp): properties
that do not need a full CPU reference, plus bit-exact parity with the oracle
on a seeded subset of pixels rendered at full spp."""
import json

import numpy as np
import pytest

import oracle_lib as O
import zig_ornament_amd as zo
from helpers import same_bits, zo_params
from zig_ornament_amd import workloads as W

pytestmark = pytest.mark.gpu


def check_subset(prod, spec, acc_g, rng_g, pixels, rng0=None):
    """The oracle renders `pixels` from their seeded RNG state, or from
    rng0[pixels] (the state an earlier render left)."""
    n = len(pixels)
    acc = np.zeros((n, 4), np.float32)
    rng = pixels.astype(np.uint32).copy() if rng0 is None else rng0[pixels].copy()
    p = O.ConstantParams.from_buffer_copy(bytes(zo_params(spec)))
    O.OracleScene(prod).render_pixels(p, spec.spp, pixels, acc, rng)
    assert same_bits(acc, acc_g[pixels])
    assert rng.tobytes() == rng_g[pixels].tobytes()


def cached_entry(tuning, ms):
    """The shipped tuning-cache entry a cache hit returned (bench.py applies
    the same one)."""
    entries = json.load(open(zo.TUNING_CACHE)).values()
    hits = [e for e in entries if e["ms"] == pytest.approx(ms, rel=1e-6)
            and all(e[k] == v for k, v in tuning.items() if k in e)]
    assert len(hits) == 1, (tuning, ms)
    return hits[0]


@pytest.mark.parametrize("name", ["C2", "C3"])
def test_full_size_render(device, name):
    """The launch bench.py times (the shipped tuning-cache entry of the
    scene class: several nodes per round; for C2, queues ordered by the
    previous render's pixel costs) at full size; the spp/2 + spp/2 render
    uses a different schedule and must give the same bits, a seeded subset
    must equal the oracle, and so must a subset of the second, cost-ordered
    render."""
    spec = W.random_spheres() if name == "C2" else W.instanced_icospheres()
    print(f"[full] {name} {spec.width}x{spec.height} {spec.spp} spp depth {spec.depth}", flush=True)
    prod = zo.Bvh(W.build_scene(spec))
    ds = zo.DeviceScene(device, prod)
    a = zo.Target(ds, spec.width, spec.height)
    tuning, ms = a.autotune(zo_params(spec), spec.spp, cache_path=zo.TUNING_CACHE)
    entry = cached_entry(tuning, ms)
    print(f"[full] {name} timed launch tuning {entry}", flush=True)
    assert a.tuning() == tuning
    if name == "C2":  # the headline launch queues pixels by their previous cost
        assert tuning["cost_order"] == 1
    a.render(zo_params(spec), spec.spp)
    acc_a, rng_a = a.accumulation(), a.rng_state()
    print(f"[full] {name} render {a.last_render_ms():.1f} ms", flush=True)
    # every pixel accumulated exactly spp samples. A sample can be NaN
    # (about 1 in 10^8): the reference's own FP32 arithmetic reaches
    # normalize(0) when refract() returns the zero vector at the
    # total-internal-reflection boundary; the oracle reproduces those pixels.
    assert np.all(acc_a[:, 3] == spec.spp)
    bad = ~np.all(np.isfinite(acc_a), axis=1)
    print(f"[full] {name} non-finite pixels: {int(bad.sum())}", flush=True)
    assert bad.mean() < 1e-5
    # determinism and progressive equivalence: spp/2 + spp/2 == spp
    b = zo.Target(ds, spec.width, spec.height)
    b.set_tuning(regroup_lanes=16, bvh_width=2, samples_per_job=0)
    b.render(zo_params(spec), spec.spp // 2)
    b.render(zo_params(spec, current_iteration=spec.spp // 2), spec.spp // 2)
    assert same_bits(b.accumulation(), acc_a)
    assert b.rng_state().tobytes() == rng_a.tobytes()
    # bit-exact against the oracle on 1500 seeded pixels at full spp, plus
    # every pixel that came out non-finite
    pixels = np.random.default_rng(0).choice(spec.width * spec.height, 1500, replace=False)
    pixels = np.unique(np.concatenate([pixels, np.nonzero(bad)[0]])).astype(np.uint32)
    check_subset(prod, spec, acc_a, rng_a, pixels)
    if tuning["cost_order"]:
        # bench.py times renders that follow another one: their queues are
        # ordered by the previous render's pixel costs
        assert np.all(a.pixel_costs() > 0)
        a.render(zo_params(spec), spec.spp)
        print(f"[full] {name} cost-ordered render {a.last_render_ms():.1f} ms", flush=True)
        acc_2, rng_2 = a.accumulation(), a.rng_state()
        assert np.all(acc_2[:, 3] == spec.spp)
        sub = np.random.default_rng(1).choice(spec.width * spec.height, 500, replace=False).astype(np.uint32)
        check_subset(prod, spec, acc_2, rng_2, np.unique(sub), rng0=rng_a)


def test_cornell_4096_spp_convergence(device):
    """relMSE against a 4096-spp image falls like 1/spp (same scene, seed and
    RNG stream: the low-spp images are prefixes of the reference)."""
    spec = W.cornell()
    prod = zo.Bvh(W.build_scene(spec))
    ds = zo.DeviceScene(device, prod)
    t = zo.Target(ds, spec.width, spec.height)
    errs = {}
    done = 0
    ref_t = zo.Target(ds, spec.width, spec.height)
    ref_t.render(zo_params(spec), 4096)
    ref = ref_t.accumulation()[:, :3] / 4096.0
    from helpers import rel_mse
    for spp in [16, 64, 256]:
        t.render(zo_params(spec, current_iteration=float(done)), spp - done)
        done = spp
        errs[spp] = rel_mse(t.accumulation()[:, :3] / spp, ref)
    print(f"[conv] relMSE vs 4096 spp: {errs}", flush=True)
    assert errs[64] < errs[16] / 2.5 and errs[256] < errs[64] / 2.5
