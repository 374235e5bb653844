"""The oracle's shading functions against what the mathematics fixes.

The GPU kernels agree with the oracle bit for bit; these tests pin the
oracle's own material scatter, texture fetch, half-float decode, sphere uv and
mesh shading normal to closed forms computed independently in float64:
- dielectric (material.wgsl:29-59): Snell's law, the mirror law, total
  internal reflection taken with no RNG draw, exactly one draw otherwise, and
  the Schlick reflectance decided against that draw (R0 = ((1-n)/(1+n))^2 at
  normal incidence);
- metal (material.wgsl:22-27): the exact mirror at fuzz 0, absorption (black)
  when the fuzzed direction points below the surface, 3 draws either way;
- lambertian (material.wgsl:9-20): normal + the normalised unit-sphere draw;
- texture fetch (material.wgsl:91-99 with the default sampler of
  webgpu.zig:812-825): nearest texel clamp(floor(u*W), 0, W-1) in exact
  arithmetic, clamp (not wrap) outside [0, 1], rgba8unorm as byte / 255,
  missing channels as 0, the `tex_id < textures_count` rule;
- half floats: all 65,536 encodings against numpy's float16;
- sphere uv (pathtracer.wgsl:121-129) at the six axis points and at random
  points, sphere normal from the world-space centre;
- mesh shading normal (pathtracer.wgsl:131-144, transform.wgsl:16-19) of a
  non-uniformly scaled, rotated instance: perpendicular to the instanced
  triangle's world-space edges, interpolated normals through the inverse
  transpose, front/back face from the shading normal.
Each check fails under the plausible mistakes named beside it.
"""
import ctypes as C
import math
import types
from fractions import Fraction

import numpy as np
import pytest

import oracle_lib as O
from zig_ornament_amd import workloads as W

f32 = C.c_float
IOR = 1.5


# ---------------------------------------------------------------------------
# helpers
# ---------------------------------------------------------------------------

def draws_after(state, k):
    """The RNG state after k draws (random.wgsl:11-18, pinned by the golden
    streams in test_oracle_rng_math.py)."""
    s = C.c_uint32(int(state))
    for _ in range(k):
        O.lib().orc_random_u32(C.byref(s))
    return s.value


def draw_of(state):
    """The float the next draw from `state` returns."""
    s = C.c_uint32(int(state))
    return O.lib().orc_random_f32(C.byref(s))


def stream_states(seed, n):
    """n consecutive RNG states and the draw each of them gives next."""
    u = np.zeros(n, np.uint32)
    f = np.zeros(n, np.float32)
    O.lib().orc_random_stream(seed, n, u.ctypes.data_as(C.c_void_p), f.ctypes.data_as(C.c_void_p))
    states = np.concatenate([[np.uint32(seed)], u[:-1]]).astype(np.uint32)
    return states, f.astype(np.float64)


def shading_scene(materials, textures=()):
    """An oracle scene holding only materials and textures (what
    material_scatter and material_emit read)."""
    m = np.zeros(len(materials), O.MATERIAL_DTYPE)
    for i, (kind, albedo, tex, fuzz, ior) in enumerate(materials):
        m[i]["type"] = {"lambertian": 0, "metal": 1, "dielectric": 2, "light": 3}[kind]
        m[i]["albedo"] = albedo
        m[i]["albedo_texture_index"] = 0xFFFFFFFF if tex is None else tex
        m[i]["fuzz"] = fuzz
        m[i]["ior"] = ior
    arrays = types.SimpleNamespace(
        tlas_nodes=np.zeros(0, O.NODE_DTYPE), blas_nodes=np.zeros(0, O.NODE_DTYPE),
        normals=np.zeros((0, 4), np.float32), normal_indices=np.zeros(0, np.uint32),
        uvs=np.zeros((0, 2), np.float32), uv_indices=np.zeros(0, np.uint32),
        transforms=np.zeros((0, 16), np.float32), materials=m, textures=list(textures))
    return O.OracleScene(arrays)


def params(textures_count=0):
    p = O.ConstantParams()
    p.textures_count = textures_count
    p.depth, p.width, p.height = 8, 2, 2
    p.ray_cast_epsilon = 0.001
    p.inverted_gamma = 1.0
    return p


def record(p, normal, front, material=0, uv=(0.0, 0.0)):
    h = O.HitRecord()
    h.p[:] = [float(x) for x in p]
    h.normal[:] = [float(x) for x in np.asarray(normal, np.float32)]
    h.uv[:] = [float(x) for x in uv]
    h.material_index = material
    h.t = 1.0
    h.front_face = 1 if front else 0
    return h


def scatter(scene, rd, rec, state, textures_count=0):
    """material_scatter; returns (goes on, attenuation, origin, direction,
    rng state after)."""
    s = C.c_uint32(int(state))
    att, so, sd = (f32 * 3)(), (f32 * 3)(), (f32 * 3)()
    p = params(textures_count)
    go = O.lib().orc_material_scatter(C.byref(scene.s), C.byref(p), O.fa(rd), C.byref(rec), C.byref(s),
                                      att, so, sd)
    return bool(go), np.array(att[:], np.float64), np.array(so[:], np.float64), np.array(sd[:], np.float64), s.value


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def incident(n, theta_deg, scale=3.7):
    """A float32 direction at angle theta to -n (scaled: the reference does
    not normalise ray directions), and its float64 unit vector."""
    n = unit(n)
    t = unit(np.cross(n, [0.3, -0.7, 0.5]))
    th = math.radians(theta_deg)
    d = np.float32(scale) * (-math.cos(th) * n + math.sin(th) * t).astype(np.float32)
    return d.astype(np.float32), unit(d.astype(np.float64))


def schlick(cos_t, eta):
    """Schlick's approximation, R0 + (1 - R0)(1 - cos)^5 with
    R0 = ((1 - eta) / (1 + eta))^2, in float64."""
    r0 = ((1.0 - eta) / (1.0 + eta)) ** 2
    return r0 + (1.0 - r0) * (1.0 - cos_t) ** 5


N_SURF = np.array(unit([0.3, 1.0, -0.2]), np.float32)
P_HIT = (1.0, 2.0, 3.0)


# ---------------------------------------------------------------------------
# dielectric
# ---------------------------------------------------------------------------

def glass():
    return shading_scene([("dielectric", (0.0, 0.0, 0.0), None, 0.0, IOR)])


def refracting_state(cos_i, eta, start=1):
    """A state whose draw exceeds the Schlick reflectance by a margin."""
    states, draws = stream_states(start, 4096)
    ok = np.nonzero(draws > schlick(cos_i, eta) + 1e-3)[0]
    return int(states[ok[0]])


@pytest.mark.parametrize("front,theta", [(True, 0.0), (True, 10.0), (True, 30.0), (True, 60.0), (True, 80.0),
                                         (False, 10.0), (False, 25.0), (False, 40.0)])
def test_dielectric_refraction_obeys_snell(front, theta):
    """Entering glass (front face) eta = 1/1.5, leaving it eta = 1.5: the
    refracted direction is a unit vector in the plane of incidence, crosses the
    surface, and sin(theta_t) = eta sin(theta_i). A swapped eta, a wrong sign
    in refract or an unnormalised incident vector fails here."""
    n = N_SURF.astype(np.float64)
    rd, u = incident(N_SURF, theta)
    eta = 1.0 / IOR if front else IOR
    cos_i = -float(np.dot(u, n))
    state = refracting_state(cos_i, eta)
    go, att, so, sd, after = scatter(glass(), rd, record(P_HIT, N_SURF, front), state)
    assert go and np.array_equal(att, [1.0, 1.0, 1.0])
    assert np.array_equal(so, np.array(P_HIT, np.float32))
    assert after == draws_after(state, 1)  # refraction possible: exactly one draw
    assert abs(np.linalg.norm(sd) - 1.0) < 2e-6
    sin_i = np.linalg.norm(np.cross(u, n))
    sin_t = np.linalg.norm(np.cross(sd, n))
    assert abs(sin_t - eta * sin_i) < 3e-6
    assert np.dot(sd, n) < 0.0  # through the surface
    if sin_i > 1e-3:
        assert abs(np.dot(sd, unit(np.cross(u, n)))) < 3e-6  # in the plane of incidence
        tan_u = u - np.dot(u, n) * n
        tan_t = sd - np.dot(sd, n) * n
        assert np.dot(tan_u, tan_t) > 0.0  # bends, does not fold back


@pytest.mark.parametrize("theta", [45.0, 60.0, 85.0])
def test_dielectric_total_internal_reflection_draws_nothing(theta):
    """Leaving glass at eta sin(theta) > 1: reflection is forced, and the
    `||` short-circuits, so the RNG is not advanced (material.wgsl:47-52).
    The direction is the mirror image: unit, same angle to the normal, and
    differing from the incident vector only along the normal."""
    n = N_SURF.astype(np.float64)
    rd, u = incident(N_SURF, theta)
    assert IOR * math.sin(math.radians(theta)) > 1.0
    for state in (0, 7, 123456789, 0xFFFFFFFF):
        go, att, _, sd, after = scatter(glass(), rd, record(P_HIT, N_SURF, False), state)
        assert go and np.array_equal(att, [1.0, 1.0, 1.0])
        assert after == state
        assert abs(np.linalg.norm(sd) - 1.0) < 2e-6
        assert abs(np.dot(sd, n) + np.dot(u, n)) < 2e-6
        assert np.linalg.norm(np.cross(sd - u, n)) < 3e-6


def test_dielectric_schlick_decides_against_one_draw():
    """Where refraction is possible the path reflects exactly when
    Schlick(cos, eta) > the draw. At normal incidence the reflectance is
    R0 = ((1 - 1.5) / (1 + 1.5))^2 = 0.04 for both faces; near grazing it
    tends to 1. Draws within 1e-5 of the reflectance are skipped (float32
    rounding of the reflectance)."""
    n = N_SURF.astype(np.float64)
    assert schlick(1.0, 1 / IOR) == pytest.approx(0.04) and schlick(1.0, IOR) == pytest.approx(0.04)
    states, draws = stream_states(99, 6000)
    counts = {True: 0, False: 0}
    for front, theta in [(True, 0.0), (False, 0.0), (True, 50.0), (True, 85.0), (False, 30.0), (False, 41.0)]:
        rd, u = incident(N_SURF, theta)
        eta = 1.0 / IOR if front else IOR
        cos_i = min(-float(np.dot(u, n)), 1.0)
        r = schlick(cos_i, eta)
        sel = np.argsort(np.abs(draws - r))[:40]  # the draws nearest the threshold, both sides
        sel = np.concatenate([sel, np.arange(0, 6000, 150)])
        for k in sel:
            if abs(draws[k] - r) < 1e-5:
                continue
            go, _, _, sd, after = scatter(glass(), rd, record(P_HIT, N_SURF, front), int(states[k]))
            assert go and after == draws_after(int(states[k]), 1)
            reflected = np.dot(sd, n) > 0.0
            assert reflected == (r > draws[k]), (front, theta, r, draws[k])
            counts[reflected] += 1
    assert counts[True] > 20 and counts[False] > 20


# ---------------------------------------------------------------------------
# metal and lambertian
# ---------------------------------------------------------------------------

ALBEDO = (0.8, 0.6, 0.2)


def in_unit_sphere(state):
    out = (f32 * 3)()
    s = C.c_uint32(int(state))
    O.lib().orc_random_in_unit_sphere(C.byref(s), out)
    return np.array(out[:], np.float64)


@pytest.mark.parametrize("theta", [0.0, 35.0, 70.0])
def test_metal_without_fuzz_is_the_exact_mirror(theta):
    """fuzz 0: the direction is reflect(normalize(d), N) = u - 2 (u.N) N of
    the normalised ray (a mirror of the unnormalised ray would be 3.7 times
    longer), after the 3 draws of random_in_unit_sphere; attenuation is the
    albedo."""
    n = N_SURF.astype(np.float64)
    rd, u = incident(N_SURF, theta)
    scene = shading_scene([("metal", ALBEDO, None, 0.0, 0.0)])
    for state in (5, 77777, 0xDEADBEEF):
        go, att, so, sd, after = scatter(scene, rd, record(P_HIT, N_SURF, True), state)
        mirror = u - 2.0 * np.dot(u, n) * n
        assert go
        assert np.max(np.abs(sd - mirror)) < 2e-6
        assert np.array_equal(att, np.array(ALBEDO, np.float32))
        assert after == draws_after(state, 3)


def test_metal_absorbs_below_the_surface():
    """A grazing ray with fuzz 1: the path ends (black: metal emits nothing)
    exactly when mirror + fuzz * in_unit_sphere points below the surface
    (material.wgsl:26); either way 3 draws are used."""
    n = N_SURF.astype(np.float64)
    rd, u = incident(N_SURF, 86.0)
    scene = shading_scene([("metal", ALBEDO, None, 1.0, 0.0)])
    mirror = u - 2.0 * np.dot(u, n) * n
    states, _ = stream_states(2024, 400)
    seen = {True: 0, False: 0}
    for state in states:
        s = in_unit_sphere(state)
        expect = mirror + s
        margin = float(np.dot(expect, n))
        go, att, _, sd, after = scatter(scene, rd, record(P_HIT, N_SURF, True), int(state))
        assert after == draws_after(int(state), 3)
        assert np.max(np.abs(sd - expect)) < 3e-6
        if abs(margin) < 1e-5:
            continue
        assert go == (margin > 0.0)
        seen[go] += 1
        if not go:
            emit = (f32 * 3)()
            rec = record(P_HIT, N_SURF, True)
            O.lib().orc_material_emit(C.byref(scene.s), C.byref(params()), C.byref(rec), emit)
            assert list(emit) == [0.0, 0.0, 0.0]
    assert seen[True] > 50 and seen[False] > 50


def test_lambertian_is_normal_plus_normalised_sphere_draw():
    """material.wgsl:9-20: N + normalize(random_in_unit_sphere), 3 draws,
    attenuation the albedo."""
    n = N_SURF.astype(np.float64)
    rd, _ = incident(N_SURF, 20.0)
    scene = shading_scene([("lambertian", ALBEDO, None, 0.0, 0.0)])
    states, _ = stream_states(31337, 64)
    for state in states:
        go, att, so, sd, after = scatter(scene, rd, record(P_HIT, N_SURF, True), int(state))
        assert go and after == draws_after(int(state), 3)
        assert np.max(np.abs((sd - n) - unit(in_unit_sphere(state)))) < 3e-6
        assert np.array_equal(att, np.array(ALBEDO, np.float32))


# ---------------------------------------------------------------------------
# textures
# ---------------------------------------------------------------------------

TW, TH = 5, 3


def texel_bytes(x, y):
    i = y * TW + x
    return (10 * i + 1, 250 - 9 * i, 7 * i + 3, 200 - i)


def rgba8_texture(comps=4):
    data = np.zeros((TH, TW, comps), np.uint8)
    for y in range(TH):
        for x in range(TW):
            data[y, x] = texel_bytes(x, y)[:comps]
    return {"data": data.reshape(-1), "width": TW, "height": TH, "num_components": comps,
            "bytes_per_component": 1, "is_hdr": False}


def ctex(t):
    keep = np.ascontiguousarray(t["data"]).view(np.uint8)
    tx = O.Texture()
    tx.width, tx.height = t["width"], t["height"]
    tx.num_components, tx.bytes_per_component = t["num_components"], t["bytes_per_component"]
    tx.is_hdr = 1 if t["is_hdr"] else 0
    tx.data = keep.ctypes.data
    return tx, keep


def sample(t, u, v):
    tx, keep = ctex(t)
    out = (f32 * 3)()
    O.lib().orc_sample_texture(C.byref(tx), O.fa([u, v]), out)
    del keep
    return list(out)


def nearest(coord, size):
    """clamp(floor(coord * size), 0, size - 1) in exact rational arithmetic on
    the float32 coordinate."""
    c = Fraction(float(np.float32(coord)))
    return min(max(math.floor(c * size), 0), size - 1)


def unorm8(b):
    """byte / 255 rounded to float32 (through float64, exact to well under
    half a float32 ulp for these quotients)."""
    return float(np.float32(int(b) / 255.0))


def coords(size):
    out = [-1.0, -0.3, -1e-30, 0.0, 1.0, 1.2, 7.0]
    for k in range(size + 1):
        c = np.float32(k / size)
        out += [float(c), float(np.nextafter(c, np.float32(2))), float(np.nextafter(c, np.float32(-1)))]
    return out


def test_texture_nearest_clamp_to_edge_unorm8():
    """Every u at k/W, k/W +- 1 ulp, below 0 and above 1 (and v likewise)
    selects texel clamp(floor(u W), 0, W - 1): clamp, not wrap (u = 1.2 reads
    the last column, a wrapping sampler the second); the value is byte/255."""
    t = rgba8_texture()
    for u in coords(TW):
        for v in coords(TH):
            x, y = nearest(u, TW), nearest(v, TH)
            expect = [unorm8(b) for b in texel_bytes(x, y)[:3]]
            assert sample(t, u, v) == expect, (u, v, x, y)


def test_unorm8_is_byte_over_255_for_every_byte():
    data = np.arange(256, dtype=np.uint8).repeat(4)
    t = {"data": data, "width": 256, "height": 1, "num_components": 4, "bytes_per_component": 1,
         "is_hdr": False}
    for b in range(256):
        got = sample(t, (b + 0.5) / 256.0, 0.5)
        assert got == [np.float32(b) / np.float32(255)] * 3
        assert abs(Fraction(got[0]) - Fraction(b, 255)) <= Fraction(1, 2**25)


@pytest.mark.parametrize("comps", [1, 2])
def test_texture_missing_channels_read_zero(comps):
    """r8unorm / rg8unorm sample as (r, 0, 0) / (r, g, 0) (WebGPU texel
    conversion fills missing colour channels with 0)."""
    t = rgba8_texture(comps)
    for (u, v) in [(0.1, 0.1), (0.95, 0.5), (0.5, 0.9)]:
        x, y = nearest(u, TW), nearest(v, TH)
        b = texel_bytes(x, y)
        expect = [unorm8(b[0]), unorm8(b[1]) if comps == 2 else 0.0, 0.0]
        assert sample(t, u, v) == expect


def test_half_float_decode_all_encodings():
    """Every binary16 encoding against numpy's float16 (0x3C00 = 1,
    0x0001 = 2^-24, 0x7BFF = 65504, signed zeros, infinities, NaNs)."""
    h = np.arange(65536, dtype=np.uint32)
    ref = h.astype(np.uint16).view(np.float16).astype(np.float32)
    got = np.array([O.lib().orc_half_to_float(int(x)) for x in h], np.float32)
    nan = np.isnan(ref)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint32), ref[~nan].view(np.uint32))
    assert O.lib().orc_half_to_float(0x3C00) == 1.0
    assert O.lib().orc_half_to_float(0x0001) == 2.0 ** -24
    assert O.lib().orc_half_to_float(0x7BFF) == 65504.0


def test_hdr_texture_reads_halves():
    vals = np.array([[0x3C00, 0x0001, 0x7BFF, 0x3555], [0xC000, 0x3800, 0x0000, 0x3C00]], np.uint16)
    t = {"data": vals.reshape(-1).view(np.uint8), "width": 2, "height": 1, "num_components": 4,
         "bytes_per_component": 2, "is_hdr": True}
    for x in range(2):
        expect = vals[x, :3].view(np.float16).astype(np.float32).tolist()
        assert sample(t, (x + 0.5) / 2, 0.5) == expect


def test_texture_index_below_textures_count_only():
    """material.wgsl:92: a material samples its texture only when
    tex_id < textures_count (the uniform, not the bound array's length);
    otherwise the albedo is the colour."""
    t = rgba8_texture()
    scene = shading_scene([("lambertian", ALBEDO, 0, 0.0, 0.0), ("light", ALBEDO, 1, 0.0, 0.0)],
                          textures=[t, t])
    rd, _ = incident(N_SURF, 10.0)
    uv = (0.45, 0.8)
    tex = [unorm8(b) for b in texel_bytes(nearest(uv[0], TW), nearest(uv[1], TH))[:3]]
    _, att, _, _, _ = scatter(scene, rd, record(P_HIT, N_SURF, True, 0, uv), 1, textures_count=1)
    assert att.tolist() == tex
    _, att, _, _, _ = scatter(scene, rd, record(P_HIT, N_SURF, True, 0, uv), 1, textures_count=0)
    assert att.tolist() == list(np.array(ALBEDO, np.float32).astype(np.float64))
    for count, expect in [(2, tex), (1, list(np.array(ALBEDO, np.float32).astype(np.float64)))]:
        emit = (f32 * 3)()
        p = params(count)
        rec = record(P_HIT, N_SURF, True, 1, uv)
        O.lib().orc_material_emit(C.byref(scene.s), C.byref(p), C.byref(rec), emit)
        assert list(emit) == expect


# ---------------------------------------------------------------------------
# sphere normal and uv, mesh shading normal: real closest hits
# ---------------------------------------------------------------------------

def closest_hit(scene, o, d):
    h = scene.trace(np.array([list(o) + list(d)], np.float32))[0]
    assert h["hit"] == 1
    hit = O.Hit()
    for k in ("hit", "t", "material_index", "node_type", "inverted_transform_id", "tri_id"):
        setattr(hit, k, int(h[k]) if k != "t" else float(h[k]))
    hit.uv[:] = [float(h["uv"][0]), float(h["uv"][1])]
    rec = O.HitRecord()
    O.lib().orc_make_hit_record(C.byref(scene.s), O.fa(o), O.fa(d), C.byref(hit), C.byref(rec))
    return hit, rec


def sphere_scene(center, radius):
    spec = W.SceneSpec("sphere", 4, 4, 1, 4, camera=((0, 0, 5), (0, 0, 0), (0, 1, 0), 1.0, 40.0, 0.0, 1.0))
    spec.materials.append(("lambertian", (0.5, 0.5, 0.5), None))
    spec.spheres.append((center, radius, 0))
    return O.OracleScene(O.build_bvh(spec))


SPHERE_C = (0.5, -2.0, 3.0)
SPHERE_R = 2.0


@pytest.mark.parametrize("axis,uv", [((1, 0, 0), (0.5, 0.5)), ((0, 0, 1), (0.25, 0.5)),
                                     ((0, 0, -1), (0.75, 0.5)), ((-1, 0, 0), (0.0, 0.5)),
                                     ((0, 1, 0), (None, 1.0)), ((0, -1, 0), (None, 0.0))])
def test_sphere_uv_at_the_axis_points(axis, uv):
    """pathtracer.wgsl:121-129: u = (atan2(-z, x) + pi) / (2 pi),
    v = acos(-y) / pi of the outward normal. At +x u = 1/2, at +z 1/4, at -z
    3/4, at -x 0 (the seam: atan2(-0, -1) = -pi); v = 0 at -y, 1 at +y, 1/2
    on the equator (u is undefined at the poles). The normal is the axis
    itself, from the world-space centre, and faces the incoming ray."""
    scene = sphere_scene(SPHERE_C, SPHERE_R)
    a = np.array(axis, np.float64)
    o = np.array(SPHERE_C) + 5.0 * a
    _, rec = closest_hit(scene, o, -a)
    assert np.max(np.abs(np.array(rec.normal[:]) - a)) < 1e-6
    assert rec.front_face == 1
    assert np.max(np.abs(np.array(rec.p[:]) - (np.array(SPHERE_C) + SPHERE_R * a))) < 1e-5
    if uv[0] is not None:
        assert rec.uv[0] == pytest.approx(uv[0], abs=2e-7)
    assert rec.uv[1] == pytest.approx(uv[1], abs=2e-7)


def test_sphere_uv_and_normal_at_random_points():
    """The same closed form at 200 random points of a sphere of radius 2 away
    from the origin (a normal taken from the object-space origin, or a
    swapped atan2 argument, fails), from outside and from inside."""
    scene = sphere_scene(SPHERE_C, SPHERE_R)
    rng = np.random.default_rng(5)
    for k in range(200):
        a = unit(rng.normal(size=3))
        if abs(a[0]) < 1e-3 and abs(a[2]) < 1e-3:
            continue
        inside = k % 4 == 3
        if inside:  # from the centre outwards: back face
            o, d = np.array(SPHERE_C), a
        else:
            o, d = np.array(SPHERE_C) + 6.0 * a, -a
        _, rec = closest_hit(scene, o, d)
        assert np.max(np.abs(np.array(rec.normal[:]) - (-a if inside else a))) < 2e-6
        assert rec.front_face == (0 if inside else 1)
        u = (math.atan2(-a[2], a[0]) + math.pi) / (2 * math.pi)
        v = math.acos(-a[1]) / math.pi
        if 1e-4 < u < 1 - 1e-4:
            assert rec.uv[0] == pytest.approx(u, abs=2e-6)
        assert rec.uv[1] == pytest.approx(v, abs=2e-6)


TRI_V = np.array([[1, 0, 0, 1], [0, 1, 0, 1], [0, 0, 1, 1]], np.float32)  # plane x + y + z = 1


def instance_matrix():
    """Row-vector (zmath) model matrix: non-uniform scale, rotation,
    translation."""
    q = unit([0.3, -0.5, 0.8, 0.4])
    m = W.mat_mul(W.mat_mul(W.mat_scaling(3.0, 0.5, 1.5), W.mat_from_quat(q)), W.mat_translation(2.0, -1.0, 4.0))
    return m.astype(np.float32)


def mesh_instance_scene(vertex_normals):
    spec = W.SceneSpec("inst", 4, 4, 1, 4, camera=((0, 0, 5), (0, 0, 0), (0, 1, 0), 1.0, 40.0, 0.0, 1.0))
    spec.materials.append(("lambertian", (0.5, 0.5, 0.5), None))
    vi = np.array([0, 1, 2], np.uint32)
    nn = np.array([list(unit(n)) + [0.0] for n in vertex_normals], np.float32)
    uvs = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]], np.float32)
    far = W.mat_translation(500.0, 500.0, 500.0)  # the mesh itself is a TLAS leaf too: out of the way
    spec.meshes.append(("mesh", TRI_V, vi, nn, vi, uvs, vi, far, 0))
    spec.instances.append((0, instance_matrix(), 0))
    return O.OracleScene(O.build_bvh(spec))


def world_triangle():
    m = instance_matrix().astype(np.float64)
    return TRI_V.astype(np.float64) @ m  # rows: world vertices (w = 1)


@pytest.mark.parametrize("bary", [(0.25, 0.35), (0.6, 0.1), (0.05, 0.9)])
def test_mesh_shading_normal_of_a_scaled_instance_is_perpendicular(bary):
    """Flat vertex normals: the shading normal of the instanced triangle is
    perpendicular to its world-space edges. Taking the normal through the
    model matrix instead of transpose(inverse) (transform.wgsl:16-19), or
    through the inverse without the transpose, breaks this under the
    non-uniform scale. Seen from the back the normal flips and front_face is
    0 (hitrecord.wgsl:10-18)."""
    scene = mesh_instance_scene([(1, 1, 1)] * 3)
    wv = world_triangle()[:, :3]
    e1, e2 = wv[1] - wv[0], wv[2] - wv[0]
    n_geo = unit(np.cross(e1, e2))
    # the object-space normal (1,1,1) maps to the side of e1 x e2 that (1,1,1) is on for this matrix
    s = 1.0 if np.linalg.det(instance_matrix()[:3, :3].astype(np.float64)) > 0 else -1.0
    n_out = s * n_geo
    target = (1 - bary[0] - bary[1]) * wv[0] + bary[0] * wv[1] + bary[1] * wv[2]
    for side in (1.0, -1.0):
        o = target + side * 4.0 * n_out
        hit, rec = closest_hit(scene, o, target - o)
        assert hit.node_type == 2
        assert hit.uv[0] == pytest.approx(bary[0], abs=1e-5) and hit.uv[1] == pytest.approx(bary[1], abs=1e-5)
        nrm = np.array(rec.normal[:], np.float64)
        assert abs(np.dot(nrm, e1)) / np.linalg.norm(e1) < 2e-6
        assert abs(np.dot(nrm, e2)) / np.linalg.norm(e2) < 2e-6
        assert np.max(np.abs(nrm - side * n_out)) < 2e-6
        assert rec.front_face == (1 if side > 0 else 0)
        # uv interpolated with the same barycentrics: uvs (0,0), (1,0), (0,1)
        assert rec.uv[0] == pytest.approx(hit.uv[0], abs=1e-6) and rec.uv[1] == pytest.approx(hit.uv[1], abs=1e-6)
        assert np.max(np.abs(np.array(rec.p[:]) - target)) < 1e-4


def test_mesh_interpolated_normal_goes_through_the_inverse_transpose():
    """Distinct vertex normals: the shading normal is
    normalize((w n0 + u n1 + v n2) A^-T) (row vectors, A the model matrix's
    3x3 part), computed here in float64 from the model matrix alone."""
    normals = [unit([1, 0.2, 0.1]), unit([0.1, 1, -0.3]), unit([0.2, 0.3, 1])]
    scene = mesh_instance_scene(normals)
    a = instance_matrix()[:3, :3].astype(np.float64)
    wv = world_triangle()[:, :3]
    n_geo = unit(np.cross(wv[1] - wv[0], wv[2] - wv[0]))
    rng = np.random.default_rng(11)
    for _ in range(40):
        u, v = rng.uniform(0.05, 0.9, 2)
        if u + v > 0.95:
            continue
        target = (1 - u - v) * wv[0] + u * wv[1] + v * wv[2]
        o = target + 3.0 * n_geo
        hit, rec = closest_hit(scene, o, target - o)
        hu, hv = hit.uv[0], hit.uv[1]
        nrm_obj = (1 - hu - hv) * normals[0] + hu * normals[1] + hv * normals[2]
        expect = unit(nrm_obj @ np.linalg.inv(a).T)
        d = target - o
        if np.dot(d, expect) > 0:
            expect = -expect
        assert np.max(np.abs(np.array(rec.normal[:]) - expect)) < 3e-6
