"""Synthetic scenes of the benchmark configurations, as plain specs.

A SceneSpec lists, in creation order, the textures, materials, spheres,
meshes and mesh instances a program would create through the Scene API
(scene.zig), plus the camera and render settings. `build_scene` replays it
through this library's Scene; the tests replay the same spec through the
oracle, so both sides build from identical inputs.

Random choices use SplitMix64 (seeded), so every scene is reproducible:
    C1  Cornell box (glfw_example/examples.zig:359-456) scaled to side 2,
        with a dielectric and a metal sphere of radius 0.3; 256x256, 16 spp, depth 8
    C2  random spheres (examples.zig:33-91): ground, a 22x22 jittered grid of
        r=0.2 spheres, three r=1 spheres; 1920x1080, 64 spp, depth 16
    C3  a 7-level icosphere (327,680 triangles) instanced 1,024 times with
        random rotation, scale U(0.5, 2) and translation in a 100^3 box,
        textured by 1024^2 RGBA8 Perlin and checker textures; 1920x1080, 32 spp, depth 8
    C4  a batch of 64 instanced-icosphere scenes, 1e3..2e6 triangles
        (log-uniform), 128^2..1024^2 pixels, 1..64 spp, depth 8
"""
import dataclasses
import math
import os

import numpy as np

MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014): the scene generators' PRNG."""

    def __init__(self, seed):
        self.s = seed & MASK64

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def float(self):
        """Uniform in [0, 1) with 24 random bits (an exact f32)."""
        return (self.next() >> 40) * (1.0 / 16777216.0)


@dataclasses.dataclass
class SceneSpec:
    name: str
    width: int
    height: int
    spp: int
    depth: int
    camera: tuple  # (lookfrom, lookat, vup, aspect, vfov, aperture, focus_dist)
    textures: list = dataclasses.field(default_factory=list)   # dicts
    materials: list = dataclasses.field(default_factory=list)  # (kind, albedo|("tex", i), param)
    spheres: list = dataclasses.field(default_factory=list)    # (center, radius, material)
    meshes: list = dataclasses.field(default_factory=list)     # ("plane", ...) | ("mesh", ...)
    instances: list = dataclasses.field(default_factory=list)  # (mesh, transform, material)
    ray_cast_epsilon: float = 0.001
    gamma: float = 1.0
    flip_y: bool = False
    seed: int = 0

    @property
    def samples(self):
        return self.width * self.height * self.spp


# ---------------------------------------------------------------------------
# row-vector (zmath-shaped) matrices built in float32
# ---------------------------------------------------------------------------

def mat_scaling(x, y, z):
    return np.diag(np.array([x, y, z, 1.0], dtype=np.float32))


def mat_translation(x, y, z):
    m = np.eye(4, dtype=np.float32)
    m[3, :3] = np.array([x, y, z], dtype=np.float32)
    return m


def mat_from_quat(q):
    x, y, z, w = [np.float32(v) for v in q]
    x2, y2, z2 = x + x, y + y, z + z
    xx, yy, zz = x * x2, y * y2, z * z2
    xy, xz, yz = x * y2, x * z2, y * z2
    wx, wy, wz = w * x2, w * y2, w * z2
    one = np.float32(1.0)
    return np.array([[(one - yy) - zz, xy + wz, xz - wy, 0],
                     [xy - wz, (one - xx) - zz, yz + wx, 0],
                     [xz + wy, yz - wx, (one - xx) - yy, 0],
                     [0, 0, 0, 1]], dtype=np.float32)


def mat_mul(a, b):
    return (np.asarray(a, np.float64) @ np.asarray(b, np.float64)).astype(np.float32)


# ---------------------------------------------------------------------------
# geometry
# ---------------------------------------------------------------------------

def icosphere(levels):
    """Unit icosphere: 20 * 4**levels triangles, 10 * 4**levels + 2 vertices."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    for _ in range(levels):
        edges = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        key = np.sort(edges, axis=1)
        uniq, inv = np.unique(key, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = len(v)
        v = np.concatenate([v, mid])
        nf = len(f)
        ab, bc, ca = inv[:nf] + base, inv[nf:2 * nf] + base, inv[2 * nf:] + base
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        f = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1),
                            np.stack([c, ca, bc], 1), np.stack([ab, bc, ca], 1)])
    return v, f


def sphere_uv(p):
    """math.zig:43-47, in double precision."""
    theta = np.arccos(-p[:, 1])
    phi = np.arctan2(-p[:, 2], p[:, 0]) + math.pi
    return np.stack([phi / (2 * math.pi), theta / math.pi], 1)


def icosphere_mesh(levels):
    v, f = icosphere(levels)
    verts = np.concatenate([v, np.ones((len(v), 1))], 1).astype(np.float32)
    normals = np.concatenate([v, np.zeros((len(v), 1))], 1).astype(np.float32)
    uvs = sphere_uv(v).astype(np.float32)
    idx = f.reshape(-1).astype(np.uint32)
    return verts, idx, normals, idx.copy(), uvs, idx.copy()


def perlin_texture(size, seed, cells=16):
    """Seeded gradient noise, 3 octaves, mapped to an earthy RGBA8 ramp."""
    rng = SplitMix64(seed)
    out = np.zeros((size, size), dtype=np.float64)
    amp, total = 1.0, 0.0
    for octave in range(3):
        n = cells * (2 ** octave)
        ang = np.array([rng.float() for _ in range((n + 1) * (n + 1))]).reshape(n + 1, n + 1)
        ang *= 2 * math.pi
        gx, gy = np.cos(ang), np.sin(ang)
        # gradients repeat at the border so the texture tiles
        gx[-1, :], gy[-1, :] = gx[0, :], gy[0, :]
        gx[:, -1], gy[:, -1] = gx[:, 0], gy[:, 0]
        c = (np.arange(size) + 0.5) * n / size
        i0 = np.floor(c).astype(int)
        fr = c - i0
        X, Y = np.meshgrid(i0, i0, indexing="xy")
        FX, FY = np.meshgrid(fr, fr, indexing="xy")

        def dotg(ix, iy, dx, dy):
            return gx[iy, ix] * dx + gy[iy, ix] * dy

        n00 = dotg(X, Y, FX, FY)
        n10 = dotg(X + 1, Y, FX - 1, FY)
        n01 = dotg(X, Y + 1, FX, FY - 1)
        n11 = dotg(X + 1, Y + 1, FX - 1, FY - 1)
        sx = FX * FX * FX * (FX * (FX * 6 - 15) + 10)
        sy = FY * FY * FY * (FY * (FY * 6 - 15) + 10)
        nx0 = n00 + sx * (n10 - n00)
        nx1 = n01 + sx * (n11 - n01)
        out += amp * (nx0 + sy * (nx1 - nx0))
        total += amp
        amp *= 0.5
    t = np.clip(0.5 + out / total, 0.0, 1.0)
    rgb = np.stack([0.2 + 0.7 * t, 0.15 + 0.55 * t * t, 0.1 + 0.3 * (1 - t)], -1)
    rgba = np.concatenate([rgb, np.ones((size, size, 1))], -1)
    return np.round(rgba * 255).astype(np.uint8)


def checker_texture(size, squares=16):
    q = np.arange(size) * squares // size
    mask = (q[None, :] + q[:, None]) % 2 == 0
    rgba = np.empty((size, size, 4), dtype=np.uint8)
    rgba[mask] = (230, 230, 220, 255)
    rgba[~mask] = (40, 90, 160, 255)
    return rgba


# ---------------------------------------------------------------------------
# the configurations
# ---------------------------------------------------------------------------

def cornell(width=256, height=256, spp=16, depth=8):
    """C1: examples.zig:359-456 scaled by 2/555, plus two r=0.3 spheres."""
    s = 2.0 / 555.0
    spec = SceneSpec("cornell", width, height, spp, depth,
                     camera=((278 * s, 278 * s, -800 * s), (278 * s, 278 * s, 0.0), (0, 1, 0),
                             width / height, 40.0, 0.0, 10.0))
    spec.materials = [("lambertian", (0.65, 0.05, 0.05), None),   # 0 red
                      ("lambertian", (0.73, 0.73, 0.73), None),   # 1 white
                      ("lambertian", (0.12, 0.45, 0.15), None),   # 2 green
                      ("light", (15.0, 15.0, 15.0), None),        # 3 light
                      ("dielectric", None, 1.5),                  # 4
                      ("metal", (0.8, 0.85, 0.88), 0.1)]          # 5

    def quad(q, u, v):
        return tuple(q[i] + 0.5 * u[i] + 0.5 * v[i] for i in range(3))

    walls = [
        (quad((555, 0, 0), (0, 555, 0), (0, 0, 555)), 555, 555, (1, 0, 0), 2),
        (quad((0, 0, 0), (0, 555, 0), (0, 0, 555)), 555, 555, (-1, 0, 0), 0),
        (quad((343, 554, 332), (-130, 0, 0), (0, 0, -105)), 130, 105, (0, -1, 0), 3),
        (quad((0, 0, 0), (555, 0, 0), (0, 0, 555)), 555, 555, (0, 1, 0), 1),
        (quad((555, 555, 555), (-555, 0, 0), (0, 0, -555)), 555, 555, (0, -1, 0), 1),
        (quad((0, 0, 555), (555, 0, 0), (0, 555, 0)), 555, 555, (0, 0, 1), 1),
    ]
    for c, a, b, n, m in walls:
        spec.meshes.append(("plane", tuple(x * s for x in c), a * s, b * s, n, m))
    spec.spheres = [((1.35, 0.3, 0.8), 0.3, 4), ((0.62, 0.3, 1.25), 0.3, 5)]
    return spec


def random_spheres(width=1920, height=1080, spp=64, depth=16, seed=1, grid=range(-11, 11)):
    """C2: examples.zig:33-91 with a 22x22 grid and SplitMix64 draws."""
    rng = SplitMix64(seed)
    spec = SceneSpec("random_spheres", width, height, spp, depth,
                     camera=((13.0, 2.0, 3.0), (0.0, 0.0, 0.0), (0, 1, 0), width / height, 20.0,
                             0.1, 10.0))
    spec.materials.append(("lambertian", (0.5, 0.5, 0.5), None))
    spec.spheres.append(((0.0, -1000.0, 0.0), 1000.0, 0))
    for a in grid:
        for b in grid:
            choose = rng.float()
            center = (a + 0.9 * rng.float(), 0.2, b + 0.9 * rng.float())
            d = math.sqrt((center[0] - 4.0) ** 2 + (center[2]) ** 2)
            if d <= 0.9:
                continue
            if choose < 0.8:
                c1 = [rng.float() for _ in range(3)]
                c2 = [rng.float() for _ in range(3)]
                spec.materials.append(("lambertian", tuple(x * y for x, y in zip(c1, c2)), None))
            elif choose < 0.95:
                alb = tuple(0.5 + 0.5 * rng.float() for _ in range(3))
                spec.materials.append(("metal", alb, 0.5 * rng.float()))
            else:
                spec.materials.append(("dielectric", None, 1.5))
            spec.spheres.append((center, 0.2, len(spec.materials) - 1))
    spec.materials.append(("dielectric", None, 1.5))
    spec.spheres.append(((0.0, 1.0, 0.0), 1.0, len(spec.materials) - 1))
    spec.materials.append(("lambertian", (0.4, 0.2, 0.1), None))
    spec.spheres.append(((-4.0, 1.0, 0.0), 1.0, len(spec.materials) - 1))
    spec.materials.append(("metal", (0.7, 0.6, 0.5), 0.0))
    spec.spheres.append(((4.0, 1.0, 0.0), 1.0, len(spec.materials) - 1))
    return spec


def instanced_icospheres(width=1920, height=1080, spp=32, depth=8, levels=7, instances=1024,
                         texture_size=1024, seed=3, box=100.0):
    """C3: one icosphere BLAS instanced in the TLAS, two textured materials."""
    rng = SplitMix64(seed)
    spec = SceneSpec("instanced_icospheres", width, height, spp, depth,
                     camera=((0.0, 0.3 * box, -1.6 * box), (0.0, 0.0, 0.0), (0, 1, 0),
                             width / height, 45.0, 0.0, 10.0))
    spec.textures = [
        dict(data=perlin_texture(texture_size, seed + 1), width=texture_size, height=texture_size,
             num_components=4, bytes_per_component=1, is_hdr=False),
        dict(data=checker_texture(texture_size), width=texture_size, height=texture_size,
             num_components=4, bytes_per_component=1, is_hdr=False),
    ]
    spec.materials = [("lambertian", ("tex", 0), None), ("lambertian", ("tex", 1), None)]
    verts, vi, normals, ni, uvs, ui = icosphere_mesh(levels)
    spec.meshes.append(("mesh", verts, vi, normals, ni, uvs, ui, np.eye(4, dtype=np.float32), 1))
    for k in range(instances):
        # uniform random rotation (Shoemake), scale U(0.5, 2), translation in the box
        u1, u2, u3 = rng.float(), rng.float(), rng.float()
        q = (math.sqrt(1 - u1) * math.sin(2 * math.pi * u2), math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
             math.sqrt(u1) * math.sin(2 * math.pi * u3), math.sqrt(u1) * math.cos(2 * math.pi * u3))
        sc = 0.5 + 1.5 * rng.float()
        tr = [box * (rng.float() - 0.5) for _ in range(3)]
        m = mat_mul(mat_mul(mat_scaling(sc, sc, sc), mat_from_quat(q)), mat_translation(*tr))
        spec.instances.append((0, m, k % 2))
    return spec


def orbit_lookfrom(spec, frame, frames):
    """C5's seeded camera path: frame `frame` of `frames` on a quarter orbit
    around the vertical axis through the scene centre, starting at the
    scene's own camera position (look-at point and up vector unchanged)."""
    lookfrom = np.array(spec.camera[0], np.float64)
    a = 0.5 * np.pi * frame / max(frames - 1, 1)
    c, s = np.cos(a), np.sin(a)
    x, y, z = lookfrom
    return (float(c * x + s * z), float(y), float(-s * x + c * z))


def instanced_scene(triangles, side, spp, depth=8, seed=5, texture_size=128):
    """One image of the batch workload: icospheres instanced until the scene
    holds about `triangles` triangles (level 3, 1,280 triangles each, below
    20,480; level 4, 5,120 each, from there), in a box whose volume grows with
    the instance count; textured Lambertian, metal and dielectric instances.
    The base mesh is a TLAS leaf too (bvh.zig:101-104): it sits at the
    origin at unit scale, inside the instance cloud."""
    rng = SplitMix64(seed)
    levels = 3 if triangles < 20480 else 4
    per = 20 * 4 ** levels
    instances = max(1, int(round(triangles / per)))
    box = 6.0 * instances ** (1.0 / 3.0)
    spec = SceneSpec(f"instanced_{triangles}_{side}px_{spp}spp", side, side, spp, depth,
                     camera=((0.0, 0.4 * box, -1.5 * box), (0.0, 0.0, 0.0), (0, 1, 0), 1.0,
                             40.0, 0.0, 10.0))
    spec.textures = [
        dict(data=perlin_texture(texture_size, seed + 1), width=texture_size,
             height=texture_size, num_components=4, bytes_per_component=1, is_hdr=False),
        dict(data=checker_texture(texture_size), width=texture_size, height=texture_size,
             num_components=4, bytes_per_component=1, is_hdr=False),
    ]
    spec.materials = [("lambertian", ("tex", 0), None), ("lambertian", ("tex", 1), None),
                      ("metal", (0.8, 0.7, 0.6), 0.2), ("dielectric", None, 1.5)]
    verts, vi, normals, ni, uvs, ui = icosphere_mesh(levels)
    spec.meshes.append(("mesh", verts, vi, normals, ni, uvs, ui, np.eye(4, dtype=np.float32), 0))
    for k in range(instances):
        u1, u2, u3 = rng.float(), rng.float(), rng.float()
        q = (math.sqrt(1 - u1) * math.sin(2 * math.pi * u2), math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
             math.sqrt(u1) * math.sin(2 * math.pi * u3), math.sqrt(u1) * math.cos(2 * math.pi * u3))
        sc = 0.5 + 1.5 * rng.float()
        tr = [box * (rng.float() - 0.5) for _ in range(3)]
        m = mat_mul(mat_mul(mat_scaling(sc, sc, sc), mat_from_quat(q)), mat_translation(*tr))
        spec.instances.append((0, m, k % 4))
    return spec


def batch_scenes(count=64, seed=4, min_tris=1e3, max_tris=2e6, sides=(128, 256, 512, 1024),
                 spps=(1, 2, 4, 8, 16, 32, 64), depth=8):
    """C4: `count` images, triangle counts log-uniform in [min_tris, max_tris]
    (instanced icospheres), square resolutions drawn from `sides`, samples
    per pixel drawn from `spps`, each with its own RNG seed."""
    rng = SplitMix64(seed)
    specs = []
    for i in range(count):
        tris = int(round(math.exp(math.log(min_tris) + rng.float() * math.log(max_tris / min_tris))))
        side = sides[int(rng.float() * len(sides))]
        spp = spps[int(rng.float() * len(spps))]
        spec = instanced_scene(tris, side, spp, depth, seed=1000 + i)
        spec.seed = i
        specs.append(spec)
    return specs


CONFIGS = {
    "C1": cornell,
    "C2": random_spheres,
    "C3": instanced_icospheres,
}


def build_scene(spec):
    """Replays a SceneSpec through the Scene API; returns the Scene."""
    from . import Scene
    s = Scene()
    tex = [s.create_texture(t["data"], t["width"], t["height"], t["num_components"],
                            t["bytes_per_component"], t["is_hdr"]) for t in spec.textures]
    mats = []
    for kind, albedo, param in spec.materials:
        a = tex[albedo[1]] if isinstance(albedo, tuple) and albedo and albedo[0] == "tex" else albedo
        if kind == "lambertian":
            mats.append(s.lambertian(a))
        elif kind == "metal":
            mats.append(s.metal(a, param))
        elif kind == "dielectric":
            mats.append(s.dielectric(param))
        else:
            mats.append(s.diffuse_light(a))
    for center, radius, m in spec.spheres:
        s.attach_sphere(s.create_sphere(center, radius, mats[m]))
    meshes = []
    for entry in spec.meshes:
        if entry[0] == "plane":
            _, c, a, b, n, m = entry
            meshes.append(s.create_plane_mesh(c, a, b, n, mats[m]))
        else:
            _, v, vi, n, ni, uv, ui, t, m = entry
            meshes.append(s.create_mesh(v, vi, n, ni, uv, ui, t, mats[m]))
        s.attach_mesh(meshes[-1])
    for mesh, t, m in spec.instances:
        s.attach_mesh_instance(s.create_mesh_instance(meshes[mesh], t, mats[m]))
    lookfrom, lookat, vup, aspect, vfov, aperture, focus = spec.camera
    s.set_camera(lookfrom, lookat, vup, aspect, vfov, aperture, focus)
    return s


def spec_params(spec, current_iteration=0.0):
    """The ConstantParams (gpu_structs.zig:61-86) of a spec's render."""
    from . import _native
    p = _native.ConstantParams()
    p.camera = build_scene_camera(spec)
    p.depth, p.width, p.height = spec.depth, spec.width, spec.height
    p.flip_y = 1 if spec.flip_y else 0
    p.inverted_gamma = 1.0 / spec.gamma
    p.ray_cast_epsilon = spec.ray_cast_epsilon
    p.textures_count = len(spec.textures)
    p.current_iteration = current_iteration
    return p


def spec_settings(spec):
    """The render settings of a spec, as stored in a scene file."""
    return dict(width=spec.width, height=spec.height, iterations=spec.spp, depth=spec.depth,
                gamma=spec.gamma, ray_cast_epsilon=spec.ray_cast_epsilon,
                flip_y=1 if spec.flip_y else 0, seed=spec.seed)


def save_spec(spec, path):
    """Writes a spec as a scene file (ornament_scene_save of the Scene the
    spec builds, with its render settings)."""
    build_scene(spec).save(path, spec_settings(spec))


_MATERIAL_KINDS = {0: "lambertian", 1: "metal", 2: "dielectric", 3: "light"}


def load_spec(path):
    """Reads a scene file (the ORNSCENE layout ornament_scene_save writes)
    into a SceneSpec whose meshes are explicit arrays; building it with
    build_scene gives the Scene ornament_scene_load gives."""
    b = open(path, "rb").read()
    at = [0]

    def take(fmt_dtype, count):
        dt = np.dtype(fmt_dtype)
        n = dt.itemsize * count
        if at[0] + n > len(b):
            raise ValueError(f"{path}: truncated scene file")
        a = np.frombuffer(b, dt, count, at[0])
        at[0] += n
        return a

    if b[:8] != b"ORNSCENE":
        raise ValueError(f"{path}: not a scene file")
    at[0] = 8
    version, flags = take("<u4", 2)
    if version != 1:
        raise ValueError(f"{path}: scene file version {version}")
    cam = take("<f4", 13)
    settings = None
    if flags & 1:
        w, h, it, depth = (int(x) for x in take("<u4", 4))
        gamma, eps = (float(x) for x in take("<f4", 2))
        flip, seed = (int(x) for x in take("<u4", 2))
        settings = dict(width=w, height=h, iterations=it, depth=depth, gamma=gamma,
                        ray_cast_epsilon=eps, flip_y=flip, seed=seed)
    n_tex, n_mat, n_sph, n_mesh, n_inst = (int(x) for x in take("<u4", 5))
    st = settings or dict(width=1, height=1, iterations=1, depth=1, gamma=1.0,
                          ray_cast_epsilon=0.001, flip_y=0, seed=0)
    camera = (tuple(float(x) for x in cam[0:3]), tuple(float(x) for x in cam[3:6]),
              tuple(float(x) for x in cam[6:9]), float(cam[9]), float(cam[10]), float(cam[11]),
              float(cam[12]))
    spec = SceneSpec(os.path.basename(path), st["width"], st["height"], st["iterations"],
                     st["depth"], camera, ray_cast_epsilon=st["ray_cast_epsilon"],
                     gamma=st["gamma"], flip_y=bool(st["flip_y"]), seed=st["seed"])
    for _ in range(n_tex):
        w, h, comps, bpc, hdr = (int(x) for x in take("<u4", 5))
        take("<f4", 1)  # gamma of the texture (not used by the kernels)
        n = int(take("<u8", 1)[0])
        data = take("u1", n).copy()
        at[0] += (-at[0]) % 4
        spec.textures.append(dict(data=data, width=w, height=h, num_components=comps,
                                  bytes_per_component=bpc, is_hdr=bool(hdr)))
    for _ in range(n_mat):
        kind, tex = (int(x) for x in take("<u4", 2))
        albedo = tuple(float(x) for x in take("<f4", 3))
        fuzz, ior = (float(x) for x in take("<f4", 2))
        a = ("tex", tex) if tex != 0xFFFFFFFF else albedo
        param = fuzz if kind == 1 else ior if kind == 2 else None
        spec.materials.append((_MATERIAL_KINDS[kind], None if kind == 2 else a, param))
    for _ in range(n_sph):
        c = tuple(float(x) for x in take("<f4", 3))
        r = float(take("<f4", 1)[0])
        m, _att = (int(x) for x in take("<u4", 2))
        spec.spheres.append((c, r, m))
    for _ in range(n_mesh):
        nv, nvi, nn, nni, nuv, nui = (int(x) for x in take("<u8", 6))
        v = take("<f4", 4 * nv).reshape(-1, 4).copy()
        vi = take("<u4", nvi).copy()
        nrm = take("<f4", 4 * nn).reshape(-1, 4).copy()
        ni = take("<u4", nni).copy()
        uv = take("<f4", 2 * nuv).reshape(-1, 2).copy()
        ui = take("<u4", nui).copy()
        t = take("<f4", 16).reshape(4, 4).copy()
        m, _att = (int(x) for x in take("<u4", 2))
        spec.meshes.append(("mesh", v, vi, nrm, ni, uv, ui, t, m))
    for _ in range(n_inst):
        mesh = int(take("<u4", 1)[0])
        t = take("<f4", 16).reshape(4, 4).copy()
        m, _att = (int(x) for x in take("<u4", 2))
        spec.instances.append((mesh, t, m))
    if at[0] != len(b):
        raise ValueError(f"{path}: trailing bytes in scene file")
    return spec, settings


def build_scene_camera(spec):
    """The device camera (gpu_structs.Camera) of a spec, computed by Scene."""
    from . import Scene
    s = Scene()
    lookfrom, lookat, vup, aspect, vfov, aperture, focus = spec.camera
    s.set_camera(lookfrom, lookat, vup, aspect, vfov, aperture, focus)
    return s.camera()
