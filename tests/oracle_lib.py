"""ctypes access to the CPU oracle (oracle/build/libornament_oracle.so).

Test infrastructure only: the checker that the HIP path is compared against.
"""
import ctypes as C
import os
import subprocess
import types

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
# ORC_LIB selects another build of the oracle (the sanitizer build of
# oracle/Makefile's sanitize target)
LIB_PATH = os.environ.get("ORC_LIB") or os.path.join(ORACLE_DIR, "build", "libornament_oracle.so")

f32, u32, P = C.c_float, C.c_uint32, C.POINTER
vp = C.c_void_p


class Camera(C.Structure):
    _fields_ = [("origin", f32 * 3), ("lens_radius", f32), ("lower_left_corner", f32 * 3),
                ("pad0", u32), ("horizontal", f32 * 3), ("pad1", u32), ("vertical", f32 * 3),
                ("pad2", u32), ("u", f32 * 3), ("pad3", u32), ("v", f32 * 3), ("pad4", u32),
                ("w", f32 * 3), ("pad5", u32)]


class ConstantParams(C.Structure):
    _fields_ = [("camera", Camera), ("depth", u32), ("width", u32), ("height", u32),
                ("flip_y", u32), ("inverted_gamma", f32), ("ray_cast_epsilon", f32),
                ("textures_count", u32), ("current_iteration", f32)]


class Texture(C.Structure):
    _fields_ = [("width", u32), ("height", u32), ("num_components", u32),
                ("bytes_per_component", u32), ("is_hdr", u32), ("pad", u32), ("data", vp)]


class HitRecord(C.Structure):
    _fields_ = [("p", f32 * 3), ("normal", f32 * 3), ("uv", f32 * 2), ("material_index", u32),
                ("t", f32), ("front_face", C.c_int32)]


class Hit(C.Structure):
    _fields_ = [("hit", u32), ("t", f32), ("material_index", u32), ("node_type", u32),
                ("inverted_transform_id", u32), ("tri_id", u32), ("uv", f32 * 2)]


class OScene(C.Structure):
    _fields_ = [("tlas", vp), ("tlas_len", u32), ("blas", vp), ("blas_len", u32),
                ("normals", vp), ("normals_len", u32), ("normal_indices", vp),
                ("normal_indices_len", u32), ("uvs", vp), ("uvs_len", u32),
                ("uv_indices", vp), ("uv_indices_len", u32), ("transforms", vp),
                ("transforms_len", u32), ("materials", vp), ("materials_len", u32),
                ("textures", vp), ("textures_len", u32)]


class Stats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("rays", "node_visits", "transform_fetches",
                                          "sphere_hits", "mesh_hits", "textured_hits",
                                          "samples")]


class BuildMaterial(C.Structure):
    _fields_ = [("albedo", f32 * 4), ("texture", C.c_int32), ("fuzz", f32), ("ior", f32),
                ("type", u32)]


class BuildSphere(C.Structure):
    _fields_ = [("transform", f32 * 16), ("aabb_min", f32 * 4), ("aabb_max", f32 * 4),
                ("material", u32)]


class BuildMesh(C.Structure):
    _fields_ = [("vertices", vp), ("vertices_len", u32), ("vertex_indices", vp),
                ("vertex_indices_len", u32), ("normals", vp), ("normals_len", u32),
                ("normal_indices", vp), ("normal_indices_len", u32), ("uvs", vp),
                ("uvs_len", u32), ("uv_indices", vp), ("uv_indices_len", u32),
                ("transform", f32 * 16), ("aabb_min", f32 * 4), ("aabb_max", f32 * 4),
                ("material", u32)]


class BuildInstance(C.Structure):
    _fields_ = [("mesh", u32), ("transform", f32 * 16), ("aabb_min", f32 * 4),
                ("aabb_max", f32 * 4), ("material", u32)]


class BuildInput(C.Structure):
    _fields_ = [("materials", vp), ("materials_len", u32), ("spheres", vp), ("spheres_len", u32),
                ("meshes", vp), ("meshes_len", u32), ("instances", vp), ("instances_len", u32)]


class BuildOutput(C.Structure):
    _fields_ = [("tlas", vp), ("tlas_len", u32), ("blas", vp), ("blas_len", u32),
                ("normals", vp), ("normals_len", u32), ("normal_indices", vp),
                ("normal_indices_len", u32), ("uvs", vp), ("uvs_len", u32),
                ("uv_indices", vp), ("uv_indices_len", u32), ("transforms", vp),
                ("transforms_len", u32), ("materials", vp), ("materials_len", u32),
                ("textures", vp), ("textures_len", u32)]


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        sig = {
            "orc_random_u32": (u32, [P(u32)]),
            "orc_random_f32": (f32, [P(u32)]),
            "orc_random_stream": (None, [u32, u32, vp, vp]),
            "orc_random_in_unit_sphere": (None, [P(u32), P(f32)]),
            "orc_random_unit_vector": (None, [P(u32), P(f32)]),
            "orc_random_in_unit_disk": (None, [P(u32), P(f32)]),
            "orc_sinf": (f32, [f32]), "orc_cosf": (f32, [f32]), "orc_powf": (f32, [f32, f32]),
            "orc_acosf": (f32, [f32]), "orc_atan2f": (f32, [f32, f32]),
            "orc_eval_math": (None, [C.c_int, vp, vp, vp, u32]),
            "orc_camera_init": (None, [P(f32), P(f32), P(f32), f32, f32, f32, f32, P(Camera)]),
            "orc_camera_get_ray": (None, [P(Camera), f32, f32, P(u32), P(f32), P(f32)]),
            "orc_trace_rays": (None, [P(OScene), f32, vp, u32, vp, C.c_int]),
            "orc_safe_invdir": (None, [P(f32), P(f32)]),
            "orc_aabb_hit": (None, [P(f32), P(f32), P(f32), P(f32), f32, f32, P(f32)]),
            "orc_triangle_hit": (f32, [P(f32), P(f32), P(f32), P(f32), P(f32), f32, f32, P(f32)]),
            "orc_sphere_hit": (f32, [P(f32), P(f32), f32, f32]),
            "orc_brute_force": (None, [P(OScene), C.c_double, vp, u32, C.c_double, vp, C.c_int]),
            "orc_render": (f32, [P(OScene), P(ConstantParams), u32, vp, vp, vp, C.c_int, P(Stats)]),
            "orc_render_pixels": (f32, [P(OScene), P(ConstantParams), u32, vp, u32, vp, vp, vp,
                                        C.c_int, P(Stats)]),
            "orc_seed_rng": (None, [vp, u32, u32]),
            "orc_make_hit_record": (None, [P(OScene), P(f32), P(f32), P(Hit), P(HitRecord)]),
            "orc_material_scatter": (C.c_int, [P(OScene), P(ConstantParams), P(f32), P(HitRecord),
                                               P(u32), P(f32), P(f32), P(f32)]),
            "orc_material_emit": (None, [P(OScene), P(ConstantParams), P(HitRecord), P(f32)]),
            "orc_sample_texture": (None, [P(Texture), P(f32), P(f32)]),
            "orc_half_to_float": (f32, [C.c_uint16]),
            "orc_mat_inverse": (C.c_int, [P(f32), P(f32)]),
            "orc_mat_mul": (None, [P(f32), P(f32), P(f32)]),
            "orc_mat_from_quat": (None, [P(f32), P(f32)]),
            "orc_vec_mul_mat": (None, [P(f32), P(f32), P(f32)]),
            "orc_rotation_between_vectors": (None, [P(f32), P(f32), P(f32)]),
            "orc_aabb_transform": (None, [P(f32), P(f32), P(f32), P(f32), P(f32)]),
            "orc_create_sphere": (None, [P(f32), f32, P(f32), P(f32), P(f32)]),
            "orc_plane_mesh_transform": (None, [P(f32), f32, f32, P(f32), P(f32)]),
            "orc_mesh_aabb": (None, [vp, vp, u32, P(f32), P(f32), P(f32), P(f32), P(f32)]),
            "orc_bvh_build": (C.c_int, [P(BuildInput), P(BuildOutput)]),
            "orc_bvh_build_stream": (C.c_int, [P(BuildInput), P(BuildOutput), P(C.c_uint64)]),
            "orc_bvh_free": (None, [P(BuildOutput)]),
            "orc_splitmix64_next": (C.c_uint64, [P(C.c_uint64)]),
            "orc_xoshiro_seed": (None, [C.c_uint64, P(C.c_uint64)]),
            "orc_xoshiro_next": (C.c_uint64, [P(C.c_uint64)]),
            "orc_xoshiro_range_at_most": (C.c_uint64, [P(C.c_uint64), C.c_uint64]),
            "orc_heap_sort_keys": (None, [vp, vp, u32]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def fa(values, n=None):
    vals = [float(x) for x in values]
    return (f32 * (n or len(vals)))(*vals)


def ptr(a):
    return a.ctypes.data_as(vp) if a is not None and a.size else None


# ---------------------------------------------------------------------------
# scenes
# ---------------------------------------------------------------------------

NODE_DTYPE = np.dtype([("lmin_or_v0", "<f4", 3), ("left_or_custom_id", "<u4"),
                       ("lmax_or_v1", "<f4", 3), ("right_or_material_index", "<u4"),
                       ("rmin_or_v2", "<f4", 3), ("node_type", "<u4"),
                       ("rmax_or_v3", "<f4", 3), ("transform_id", "<u4")])
MATERIAL_DTYPE = np.dtype([("albedo", "<f4", 3), ("albedo_texture_index", "<u4"),
                           ("fuzz", "<f4"), ("ior", "<f4"), ("type", "<u4"), ("padding", "<u4")])
HIT_DTYPE = np.dtype([("hit", "<u4"), ("t", "<f4"), ("material_index", "<u4"),
                      ("node_type", "<u4"), ("inverted_transform_id", "<u4"), ("tri_id", "<u4"),
                      ("uv", "<f4", 2)])


class OracleScene:
    """The bound buffers of pathtracer.wgsl over numpy arrays (an object with
    the Bvh array attributes: tlas_nodes, blas_nodes, normals, ...)."""

    def __init__(self, arrays):
        self.keep = []

        def a(x, dtype):
            x = np.ascontiguousarray(x, dtype=dtype)
            self.keep.append(x)
            return x

        s = OScene()
        t = a(arrays.tlas_nodes, NODE_DTYPE)
        s.tlas, s.tlas_len = ptr(t), len(t)
        b = a(arrays.blas_nodes, NODE_DTYPE)
        s.blas, s.blas_len = ptr(b), len(b)
        n = a(arrays.normals, np.float32)
        s.normals, s.normals_len = ptr(n), len(n)
        ni = a(arrays.normal_indices, np.uint32)
        s.normal_indices, s.normal_indices_len = ptr(ni), len(ni)
        uv = a(arrays.uvs, np.float32)
        s.uvs, s.uvs_len = ptr(uv), len(uv)
        ui = a(arrays.uv_indices, np.uint32)
        s.uv_indices, s.uv_indices_len = ptr(ui), len(ui)
        tr = a(arrays.transforms, np.float32)
        s.transforms, s.transforms_len = ptr(tr), len(tr)
        m = a(arrays.materials, MATERIAL_DTYPE)
        s.materials, s.materials_len = ptr(m), len(m)
        tex = (Texture * max(1, len(arrays.textures)))()
        for i, tx in enumerate(arrays.textures):
            d = a(tx["data"], np.uint8)
            tex[i].width, tex[i].height = tx["width"], tx["height"]
            tex[i].num_components = tx["num_components"]
            tex[i].bytes_per_component = tx["bytes_per_component"]
            tex[i].is_hdr = 1 if tx["is_hdr"] else 0
            tex[i].data = d.ctypes.data
        self.keep.append(tex)
        s.textures = C.addressof(tex) if arrays.textures else None
        s.textures_len = len(arrays.textures)
        self.s = s

    def trace(self, rays, t_min=0.001, nthreads=0):
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
        out = np.zeros(len(r), dtype=HIT_DTYPE)
        lib().orc_trace_rays(C.byref(self.s), t_min, ptr(r), len(r), ptr(out), nthreads)
        return out

    def brute_force(self, rays, t_min=0.001, rel_eps=1e-5, nthreads=0):
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
        out = np.zeros((len(r), 8), dtype=np.float64)
        lib().orc_brute_force(C.byref(self.s), t_min, ptr(r), len(r), rel_eps, ptr(out), nthreads)
        return out

    def render(self, params, iterations, acc, rng, nthreads=0, stats=None):
        """One render() on every pixel; acc (N,4) f32 and rng (N,) u32 update in place."""
        n = params.width * params.height
        fb = np.zeros((n, 4), dtype=np.float32)
        assert acc.shape == (n, 4) and acc.dtype == np.float32 and acc.flags.c_contiguous
        assert rng.shape == (n,) and rng.dtype == np.uint32 and rng.flags.c_contiguous
        cur = lib().orc_render(C.byref(self.s), C.byref(params), iterations, ptr(acc), ptr(rng),
                               ptr(fb), nthreads, C.byref(stats) if stats is not None else None)
        return fb, cur

    def render_pixels(self, params, iterations, pixels, acc, rng, nthreads=0, stats=None):
        pixels = np.ascontiguousarray(pixels, dtype=np.uint32)
        fb = np.zeros((len(pixels), 4), dtype=np.float32)
        cur = lib().orc_render_pixels(C.byref(self.s), C.byref(params), iterations, ptr(pixels),
                                      len(pixels), ptr(acc), ptr(rng), ptr(fb), nthreads,
                                      C.byref(stats) if stats is not None else None)
        return fb, cur


def seed_rng(n, seed=0):
    out = np.zeros(n, dtype=np.uint32)
    lib().orc_seed_rng(ptr(out), n, seed)
    return out


def camera(lookfrom, lookat, vup, aspect, vfov, aperture, focus):
    c = Camera()
    lib().orc_camera_init(fa(lookfrom), fa(lookat), fa(vup), aspect, vfov, aperture, focus,
                          C.byref(c))
    return c


def constant_params(spec, current_iteration=0.0, textures_count=None):
    p = ConstantParams()
    lookfrom, lookat, vup, aspect, vfov, aperture, focus = spec.camera
    p.camera = camera(lookfrom, lookat, vup, aspect, vfov, aperture, focus)
    p.depth, p.width, p.height = spec.depth, spec.width, spec.height
    p.flip_y = 1 if spec.flip_y else 0
    p.inverted_gamma = 1.0 / spec.gamma
    p.ray_cast_epsilon = spec.ray_cast_epsilon
    p.textures_count = len(spec.textures) if textures_count is None else textures_count
    p.current_iteration = current_iteration
    return p


# ---------------------------------------------------------------------------
# BVH build from a SceneSpec, through the oracle's own factories
# ---------------------------------------------------------------------------

PLANE_VERTS = np.array([[-0.5, 0.0, -0.5, 1.0], [-0.5, 0.0, 0.5, 1.0], [0.5, 0.0, 0.5, 1.0],
                        [0.5, 0.0, -0.5, 1.0]], dtype=np.float32)
PLANE_IDX = np.array([3, 1, 0, 2, 1, 3], dtype=np.uint32)
PLANE_NORMALS = np.array([[0, 1, 0, 0]] * 4, dtype=np.float32)
PLANE_UVS = np.full((4, 2), 0.5, dtype=np.float32)


def build_bvh(spec, prng=None):
    """Oracle restatement of Scene factories + Bvh.init over a SceneSpec.
    prng: a (c_uint64 * 4) Xoshiro256 state to draw the split axes from and
    leave advanced (one generator across builds, bvh.zig:12-13); None seeds
    the build with 1244."""
    L = lib()
    keep = []
    mats = (BuildMaterial * max(1, len(spec.materials)))()
    type_of = {"lambertian": 0, "metal": 1, "dielectric": 2, "light": 3}
    for i, (kind, albedo, param) in enumerate(spec.materials):
        m = mats[i]
        m.type = type_of[kind]
        m.texture = -1
        if isinstance(albedo, tuple) and albedo and albedo[0] == "tex":
            m.texture = albedo[1]
        elif albedo is not None:
            m.albedo = fa(list(albedo) + [0.0] * (4 - len(albedo)), 4)
        m.fuzz = param if kind == "metal" else 0.0
        m.ior = param if kind == "dielectric" else 0.0
    spheres = (BuildSphere * max(1, len(spec.spheres)))()
    for i, (c, r, m) in enumerate(spec.spheres):
        sp = spheres[i]
        L.orc_create_sphere(fa(list(c) + [1.0]), r, sp.transform, sp.aabb_min, sp.aabb_max)
        sp.material = m
    meshes = (BuildMesh * max(1, len(spec.meshes)))()
    nt_boxes = []
    for i, entry in enumerate(spec.meshes):
        bm = meshes[i]
        if entry[0] == "plane":
            _, c, a, b, n, m = entry
            L.orc_plane_mesh_transform(fa(list(c) + [1.0]), a, b, fa(list(n) + [0.0]), bm.transform)
            v, vi, nn, ni, uv, ui = PLANE_VERTS, PLANE_IDX, PLANE_NORMALS, PLANE_IDX, PLANE_UVS, PLANE_IDX
        else:
            _, v, vi, nn, ni, uv, ui, t, m = entry
            bm.transform = fa(np.asarray(t, np.float32).reshape(16))
            if uv is None or len(uv) == 0:
                uv = np.full((len(v), 2), 0.5, dtype=np.float32)
                ui = vi
        arrs = [np.ascontiguousarray(x, dt) for x, dt in
                ((v, np.float32), (vi, np.uint32), (nn, np.float32), (ni, np.uint32),
                 (uv, np.float32), (ui, np.uint32))]
        keep.extend(arrs)
        v, vi, nn, ni, uv, ui = arrs
        bm.vertices, bm.vertices_len = ptr(v), len(v)
        bm.vertex_indices, bm.vertex_indices_len = ptr(vi), len(vi)
        bm.normals, bm.normals_len = ptr(nn), len(nn)
        bm.normal_indices, bm.normal_indices_len = ptr(ni), len(ni)
        bm.uvs, bm.uvs_len = ptr(uv), len(uv)
        bm.uv_indices, bm.uv_indices_len = ptr(ui), len(ui)
        ntmin, ntmax = (f32 * 4)(), (f32 * 4)()
        L.orc_mesh_aabb(ptr(v), ptr(vi), len(vi), bm.transform, ntmin, ntmax, bm.aabb_min,
                        bm.aabb_max)
        nt_boxes.append((ntmin, ntmax))
        bm.material = m
    inst = (BuildInstance * max(1, len(spec.instances)))()
    for i, (mesh, t, m) in enumerate(spec.instances):
        bi = inst[i]
        bi.mesh = mesh
        bi.transform = fa(np.asarray(t, np.float32).reshape(16))
        L.orc_aabb_transform(bi.transform, nt_boxes[mesh][0], nt_boxes[mesh][1], bi.aabb_min,
                             bi.aabb_max)
        bi.material = m
    bi_ = BuildInput()
    bi_.materials, bi_.materials_len = C.addressof(mats), len(spec.materials)
    bi_.spheres, bi_.spheres_len = C.addressof(spheres), len(spec.spheres)
    bi_.meshes, bi_.meshes_len = C.addressof(meshes), len(spec.meshes)
    bi_.instances, bi_.instances_len = C.addressof(inst), len(spec.instances)
    out = BuildOutput()
    if prng is None:
        rc = L.orc_bvh_build(C.byref(bi_), C.byref(out))
    else:
        rc = L.orc_bvh_build_stream(C.byref(bi_), C.byref(out), prng)
    if rc != 0:
        raise RuntimeError(f"orc_bvh_build failed: {rc}")
    try:
        def arr(p, n, dtype, tail=()):
            dtype = np.dtype(dtype)
            cnt = n * (int(np.prod(tail)) if tail else 1)
            if n == 0:
                return np.zeros((0,) + tail, dtype)
            raw = bytes((C.c_char * (cnt * dtype.itemsize)).from_address(p))
            return np.frombuffer(raw, dtype).reshape((n,) + tail).copy()

        res = types.SimpleNamespace(
            tlas_nodes=arr(out.tlas, out.tlas_len, NODE_DTYPE),
            blas_nodes=arr(out.blas, out.blas_len, NODE_DTYPE),
            normals=arr(out.normals, out.normals_len, np.float32, (4,)),
            normal_indices=arr(out.normal_indices, out.normal_indices_len, np.uint32),
            uvs=arr(out.uvs, out.uvs_len, np.float32, (2,)),
            uv_indices=arr(out.uv_indices, out.uv_indices_len, np.uint32),
            transforms=arr(out.transforms, out.transforms_len, np.float32, (16,)),
            materials=arr(out.materials, out.materials_len, MATERIAL_DTYPE),
            texture_sources=list(arr(out.textures, out.textures_len, np.uint32)),
        )
        res.textures = [spec.textures[i] for i in res.texture_sources]
        return res
    finally:
        L.orc_bvh_free(C.byref(out))
