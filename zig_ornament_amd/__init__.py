"""zig_ornament_amd -- the ornament path tracer's hot path on MI355X (gfx950).

Python face of the C-ABI in include/ornament.h and include/ornament_hip.h,
named after the reference's public API (src/ornament.zig):

    Scene            scene.zig factories + camera (camera.zig)
    Bvh              bvh.zig: TLAS/BLAS and attribute arrays, as numpy arrays
    PathTracer       the backend interface {init, state, scene, setResolution,
                     render, getFrameBuffer, deinit} (wgpu_backend/path_tracer.zig)
    Device, DeviceScene, Target
                     the thin device ABI (orn_device_open, orn_scene_upload,
                     orn_target_create, orn_render, orn_trace_rays,
                     orn_target_autotune)
    Batch            many images of many scenes in one launch (orn_batch_*)

Everything computes in the HIP library; nothing here falls back to Python.
"""
import ctypes as C
import os

import numpy as np

from . import _native as N
from ._native import OrnamentError, check, lib

__all__ = ["Scene", "Bvh", "PathTracer", "Device", "DeviceScene", "Target", "Batch",
           "OrnamentError", "MaterialType", "NODE_DTYPE", "MATERIAL_DTYPE", "HIT_DTYPE",
           "scene_desc_arrays", "lib", "TUNING_CACHE", "write_pfm", "read_pfm", "write_exr"]

# Autotuner results measured on MI355X (gfx950) for the benchmark scene
# classes; orn_target_autotune / orn_batch_autotune read and extend it.
TUNING_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_cache.json")


class MaterialType:  # material.zig:4-9
    Lambertian = 0
    Metal = 1
    Dielectric = 2
    DiffuseLight = 3


NODE_DTYPE = np.dtype([("lmin_or_v0", "<f4", 3), ("left_or_custom_id", "<u4"),
                       ("lmax_or_v1", "<f4", 3), ("right_or_material_index", "<u4"),
                       ("rmin_or_v2", "<f4", 3), ("node_type", "<u4"),
                       ("rmax_or_v3", "<f4", 3), ("transform_id", "<u4")])
MATERIAL_DTYPE = np.dtype([("albedo", "<f4", 3), ("albedo_texture_index", "<u4"),
                           ("fuzz", "<f4"), ("ior", "<f4"), ("type", "<u4"), ("padding", "<u4")])
HIT_DTYPE = np.dtype([("hit", "<u4"), ("t", "<f4"), ("material_index", "<u4"),
                      ("node_type", "<u4"), ("inverted_transform_id", "<u4"), ("tri_id", "<u4"),
                      ("uv", "<f4", 2)])
assert NODE_DTYPE.itemsize == 64 and MATERIAL_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 32


def _f4(v, w=0.0):
    a = [float(x) for x in v]
    if len(a) == 3:
        a.append(w)
    return (C.c_float * 4)(*a)


def _mat(m):
    a = np.ascontiguousarray(np.asarray(m, dtype=np.float32).reshape(16))
    return (C.c_float * 16)(*a.tolist())


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None and a.size else None


class Scene:
    """ornament.Scene (scene.zig:13-380). Objects are handles owned by the scene."""

    def __init__(self, _handle=None):
        if _handle is None:
            h = C.c_void_p()
            check(lib().ornament_scene_init(C.byref(h)), "Scene.init")
        else:
            h = _handle
        self._h = h
        self._owned = True
        self._keep = []

    # scene files -----------------------------------------------------------
    def save(self, path, settings=None):
        """ornament_scene_save: the scene (and optional render settings, a
        dict of RenderSettings fields) to an ORNSCENE file."""
        rs = None
        if settings is not None:
            rs = N.RenderSettings(**settings)
        check(lib().ornament_scene_save(self._h, C.byref(rs) if rs is not None else None,
                                        os.fsencode(path)), "ornament_scene_save")

    @classmethod
    def load(cls, path):
        """ornament_scene_load: (Scene, settings dict or None)."""
        h, rs, has = C.c_void_p(), N.RenderSettings(), C.c_int()
        check(lib().ornament_scene_load(os.fsencode(path), C.byref(h), C.byref(rs), C.byref(has)),
              "ornament_scene_load")
        settings = {f: getattr(rs, f) for f, _ in N.RenderSettings._fields_} if has.value else None
        return cls(_handle=h), settings

    @property
    def handle(self):
        return self._h

    def deinit(self):
        if self._owned and self._h:
            lib().ornament_scene_deinit(self._h)
        self._h = None

    def __del__(self):
        try:
            self.deinit()
        except Exception:
            pass

    def _out(self, fn, *args, where):
        h = C.c_void_p()
        check(fn(self._h, *args, C.byref(h)), where)
        return h

    # materials -------------------------------------------------------------
    def lambertian(self, albedo):
        if isinstance(albedo, C.c_void_p):
            return self._out(lib().ornament_scene_lambertian_texture, albedo, where="lambertian")
        return self._out(lib().ornament_scene_lambertian, _f4(albedo), where="lambertian")

    def metal(self, albedo, fuzz):
        if isinstance(albedo, C.c_void_p):
            return self._out(lib().ornament_scene_metal_texture, albedo, float(fuzz), where="metal")
        return self._out(lib().ornament_scene_metal, _f4(albedo), float(fuzz), where="metal")

    def dielectric(self, ior):
        return self._out(lib().ornament_scene_dielectric, float(ior), where="dielectric")

    def diffuse_light(self, albedo):
        if isinstance(albedo, C.c_void_p):
            return self._out(lib().ornament_scene_diffuse_light_texture, albedo,
                             where="diffuseLight")
        return self._out(lib().ornament_scene_diffuse_light, _f4(albedo), where="diffuseLight")

    # shapes ----------------------------------------------------------------
    def create_sphere(self, center, radius, material):
        return self._out(lib().ornament_scene_create_sphere, _f4(center, 1.0), float(radius),
                         material, where="createSphere")

    def create_sphere_mesh(self, center, radius, material):
        return self._out(lib().ornament_scene_create_sphere_mesh, _f4(center, 1.0), float(radius),
                         material, where="createSphereMesh")

    def create_plane_mesh(self, center, side1, side2, normal, material):
        return self._out(lib().ornament_scene_create_plane_mesh, _f4(center, 1.0), float(side1),
                         float(side2), _f4(normal, 0.0), material, where="createPlaneMesh")

    def create_mesh(self, vertices, vertex_indices, normals, normal_indices, uvs, uv_indices,
                    transform, material):
        v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 4)
        vi = np.ascontiguousarray(vertex_indices, dtype=np.uint32).reshape(-1)
        n = np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 4)
        ni = np.ascontiguousarray(normal_indices, dtype=np.uint32).reshape(-1)
        uv = np.ascontiguousarray(uvs if uvs is not None else np.zeros((0, 2)),
                                  dtype=np.float32).reshape(-1, 2)
        ui = np.ascontiguousarray(uv_indices if uv_indices is not None else np.zeros(0),
                                  dtype=np.uint32).reshape(-1)
        return self._out(lib().ornament_scene_create_mesh, _ptr(v), len(v), _ptr(vi), len(vi),
                         _ptr(n), len(n), _ptr(ni), len(ni), _ptr(uv), len(uv), _ptr(ui), len(ui),
                         _mat(transform), material, where="createMesh")

    def create_mesh_instance(self, mesh, transform, material):
        return self._out(lib().ornament_scene_create_mesh_instance, mesh, _mat(transform),
                         material, where="createMeshInstance")

    def create_texture(self, data, width, height, num_components, bytes_per_component=1,
                       is_hdr=False, gamma=1.0):
        d = np.ascontiguousarray(np.frombuffer(np.ascontiguousarray(data).tobytes(), np.uint8))
        return self._out(lib().ornament_scene_create_texture, _ptr(d), d.size, int(width),
                         int(height), int(num_components), int(bytes_per_component),
                         1 if is_hdr else 0, float(gamma), where="createTexture")

    def attach_sphere(self, s):
        check(lib().ornament_scene_attach_sphere(self._h, s), "attachSphere")

    def attach_mesh(self, m):
        check(lib().ornament_scene_attach_mesh(self._h, m), "attachMesh")

    def attach_mesh_instance(self, m):
        check(lib().ornament_scene_attach_mesh_instance(self._h, m), "attachMeshInstance")

    def counts(self):
        c = (C.c_uint32 * 5)()
        check(lib().ornament_scene_counts(self._h, c), "counts")
        return dict(zip(["spheres", "meshes", "mesh_instances", "materials", "textures"], list(c)))

    # camera ----------------------------------------------------------------
    def set_camera(self, lookfrom, lookat, vup, aspect_ratio, vfov, aperture, focus_dist):
        check(lib().ornament_scene_set_camera(self._h, _f4(lookfrom, 1.0), _f4(lookat, 1.0),
                                              _f4(vup, 0.0), float(aspect_ratio), float(vfov),
                                              float(aperture), float(focus_dist)), "Camera.init")

    def set_look_at(self, lookfrom, lookat, vup):
        check(lib().ornament_camera_set_look_at(self._h, _f4(lookfrom, 1.0), _f4(lookat, 1.0),
                                                _f4(vup, 0.0)), "Camera.setLookAt")

    def set_aspect_ratio(self, aspect_ratio):
        check(lib().ornament_camera_set_aspect_ratio(self._h, float(aspect_ratio)),
              "Camera.setAspectRatio")

    def camera(self):
        c = N.Camera()
        check(lib().ornament_camera_get(self._h, C.byref(c)), "Camera.from")
        return c

    def camera_dirty(self):
        d = C.c_int()
        check(lib().ornament_camera_get_dirty(self._h, C.byref(d)), "camera.dirty")
        return bool(d.value)


def _array(ptr, n, dtype, shape_tail=()):
    dtype = np.dtype(dtype)
    count = int(n) * int(np.prod(shape_tail)) if shape_tail else int(n)
    if n == 0 or not ptr:
        return np.zeros((0,) + tuple(shape_tail), dtype=dtype)
    buf = (C.c_char * (count * dtype.itemsize)).from_address(ptr)
    return np.frombuffer(bytes(buf), dtype=dtype).reshape((int(n),) + tuple(shape_tail)).copy()


def bvh_shared_generator(shared):
    """ornament_bvh_shared_generator: True makes the builds that follow share
    one split-axis generator, restarted from seed 1244, as the reference's
    process-wide one (bvh.zig:12-13); False (the default) seeds every build."""
    check(lib().ornament_bvh_shared_generator(1 if shared else 0), "ornament_bvh_shared_generator")


class Bvh:
    """Bvh.init (bvh.zig:33-78): builds and copies out the device arrays."""

    def __init__(self, scene, row_major_transforms=False):
        h = C.c_void_p()
        check(lib().ornament_bvh_init(scene.handle, 1 if row_major_transforms else 0,
                                      C.byref(h)), "Bvh.init")
        try:
            d = N.SceneDesc()
            check(lib().ornament_bvh_arrays(h, C.byref(d)), "Bvh.arrays")
            self.tlas_nodes = _array(d.tlas_nodes, d.tlas_nodes_len, NODE_DTYPE)
            self.blas_nodes = _array(d.blas_nodes, d.blas_nodes_len, NODE_DTYPE)
            self.normals = _array(d.normals, d.normals_len, np.float32, (4,))
            self.normal_indices = _array(d.normal_indices, d.normal_indices_len, np.uint32)
            self.uvs = _array(d.uvs, d.uvs_len, np.float32, (2,))
            self.uv_indices = _array(d.uv_indices, d.uv_indices_len, np.uint32)
            self.transforms = _array(d.transforms, d.transforms_len, np.float32, (16,))
            self.materials = _array(d.materials, d.materials_len, MATERIAL_DTYPE)
            self.textures = []
            src = (C.c_uint32 * max(1, d.textures_len))()
            check(lib().ornament_bvh_texture_sources(h, src, d.textures_len), "texture sources")
            self.texture_sources = [src[i] for i in range(d.textures_len)]
            descs = (N.TextureDesc * max(1, d.textures_len)).from_address(d.textures) \
                if d.textures_len else []
            for i in range(d.textures_len):
                t = descs[i]
                nbytes = t.width * t.height * t.num_components * t.bytes_per_component
                data = np.frombuffer(bytes((C.c_char * nbytes).from_address(t.data)), np.uint8)
                self.textures.append(dict(width=t.width, height=t.height,
                                          num_components=t.num_components,
                                          bytes_per_component=t.bytes_per_component,
                                          is_hdr=bool(t.is_hdr), data=data.copy()))
        finally:
            lib().ornament_bvh_deinit(h)


def scene_desc_arrays(arrays):
    """An orn_scene_desc over numpy arrays (kept alive by the returned tuple)."""
    keep = []

    def p(a, dtype):
        a = np.ascontiguousarray(a, dtype=dtype)
        keep.append(a)
        return (a.ctypes.data if a.size else None), a

    d = N.SceneDesc()
    d.tlas_nodes, a = p(arrays.tlas_nodes, NODE_DTYPE)
    d.tlas_nodes_len = len(a)
    d.blas_nodes, a = p(arrays.blas_nodes, NODE_DTYPE)
    d.blas_nodes_len = len(a)
    d.normals, a = p(arrays.normals, np.float32)
    d.normals_len = len(a)
    d.normal_indices, a = p(arrays.normal_indices, np.uint32)
    d.normal_indices_len = len(a)
    d.uvs, a = p(arrays.uvs, np.float32)
    d.uvs_len = len(a)
    d.uv_indices, a = p(arrays.uv_indices, np.uint32)
    d.uv_indices_len = len(a)
    d.transforms, a = p(arrays.transforms, np.float32)
    d.transforms_len = len(a)
    d.materials, a = p(arrays.materials, MATERIAL_DTYPE)
    d.materials_len = len(a)
    tex = (N.TextureDesc * max(1, len(arrays.textures)))()
    for i, t in enumerate(arrays.textures):
        data = np.ascontiguousarray(t["data"], dtype=np.uint8)
        keep.append(data)
        tex[i].width, tex[i].height = t["width"], t["height"]
        tex[i].num_components = t["num_components"]
        tex[i].bytes_per_component = t["bytes_per_component"]
        tex[i].is_hdr = 1 if t["is_hdr"] else 0
        tex[i].data = data.ctypes.data
    keep.append(tex)
    d.textures = C.addressof(tex) if arrays.textures else None
    d.textures_len = len(arrays.textures)
    return d, keep


def write_pfm(path, rgba, width, height):
    """ornament_image_write_pfm: float RGBA pixels (row 0 = bottom) as PFM."""
    a = np.ascontiguousarray(rgba, np.float32).reshape(-1, 4)
    if len(a) != width * height:
        raise ValueError("rgba must hold width * height pixels")
    check(lib().ornament_image_write_pfm(os.fsencode(path), _ptr(a), int(width), int(height)),
          "ornament_image_write_pfm")


def read_pfm(path):
    """ornament_image_read_pfm: (rgba (h*w, 4) float32 with row 0 = bottom, width, height)."""
    w, h = C.c_uint32(), C.c_uint32()
    p = os.fsencode(path)
    check(lib().ornament_image_read_pfm(p, None, 0, C.byref(w), C.byref(h)), "ornament_image_read_pfm")
    out = np.empty((w.value * h.value, 4), np.float32)
    check(lib().ornament_image_read_pfm(p, _ptr(out), len(out), C.byref(w), C.byref(h)),
          "ornament_image_read_pfm")
    return out, w.value, h.value


def write_exr(path, rgba, width, height):
    """ornament_image_write_exr: float RGBA pixels (row 0 = bottom) as OpenEXR."""
    a = np.ascontiguousarray(rgba, np.float32).reshape(-1, 4)
    if len(a) != width * height:
        raise ValueError("rgba must hold width * height pixels")
    check(lib().ornament_image_write_exr(os.fsencode(path), _ptr(a), int(width), int(height)),
          "ornament_image_write_exr")


class Device:
    """orn_device_open: one HIP device and its stream."""

    def __init__(self, device=-1):
        h = C.c_void_p()
        check(lib().orn_device_open(int(device), C.byref(h)), "orn_device_open")
        self._h = h

    @staticmethod
    def count():
        n = C.c_int()
        check(lib().orn_device_count(C.byref(n)), "orn_device_count")
        return n.value

    @property
    def handle(self):
        return self._h

    def info(self):
        buf = C.create_string_buffer(256)
        cu = C.c_int()
        check(lib().orn_device_info(self._h, buf, 256, C.byref(cu)), "orn_device_info")
        return buf.value.decode(), cu.value

    def eval_math(self, op, a, b=None):
        a = np.ascontiguousarray(a, dtype=np.float32)
        bb = np.ascontiguousarray(b if b is not None else np.zeros_like(a), dtype=np.float32)
        out = np.empty_like(a)
        check(lib().orn_eval_math(self._h, int(op), _ptr(a), _ptr(bb), _ptr(out), a.size),
              "orn_eval_math")
        return out

    def calibration_copy(self, nbytes):
        """GPU milliseconds of a plain copy kernel moving `nbytes`."""
        ms = C.c_float()
        check(lib().orn_calibration_copy(self._h, int(nbytes), C.byref(ms)), "orn_calibration_copy")
        return ms.value

    def close(self):
        if self._h:
            lib().orn_device_close(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceScene:
    """orn_scene_upload of Bvh arrays."""

    def __init__(self, device, arrays):
        d, keep = scene_desc_arrays(arrays)
        h = C.c_void_p()
        check(lib().orn_scene_upload(device.handle, C.byref(d), C.byref(h)), "orn_scene_upload")
        self._h = h
        self.device = device

    @property
    def handle(self):
        return self._h

    def stack_depth(self):
        v = C.c_uint32()
        check(lib().orn_scene_stack_depth(self._h, C.byref(v)), "orn_scene_stack_depth")
        return v.value

    def set_trace_width(self, width):
        check(lib().orn_scene_set_trace_width(self._h, int(width)), "orn_scene_set_trace_width")

    def trace(self, rays, t_min=0.001, packet=False):
        """Closest hits (orn_trace_rays; with packet, orn_trace_rays_packet)."""
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
        out = np.zeros(len(r), dtype=HIT_DTYPE)
        fn = lib().orn_trace_rays_packet if packet else lib().orn_trace_rays
        check(fn(self._h, float(t_min), _ptr(r), len(r), _ptr(out)),
              "orn_trace_rays_packet" if packet else "orn_trace_rays")
        return out

    def destroy(self):
        if self._h:
            lib().orn_scene_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def _tuning_struct(regroup_lanes=48, workgroups_per_cu=0, bvh_width=2, leaf_regroup_lanes=16,
                   samples_per_job=16, split_percent=0, node_steps=4, workgroup_pool=0, cost_order=0):
    t = N.Tuning()
    t.regroup_lanes = int(regroup_lanes)
    t.workgroups_per_cu = int(workgroups_per_cu)
    t.bvh_width = int(bvh_width)
    t.leaf_regroup_lanes = int(leaf_regroup_lanes)
    t.samples_per_job = int(samples_per_job)
    t.split_percent = int(split_percent)
    t.node_steps = int(node_steps)
    t.workgroup_pool = int(workgroup_pool)
    t.cost_order = int(cost_order)
    return t


def _tuning_dict(t):
    return {"regroup_lanes": t.regroup_lanes, "workgroups_per_cu": t.workgroups_per_cu,
            "bvh_width": t.bvh_width, "leaf_regroup_lanes": t.leaf_regroup_lanes,
            "samples_per_job": t.samples_per_job, "split_percent": t.split_percent,
            "node_steps": t.node_steps, "workgroup_pool": t.workgroup_pool, "cost_order": t.cost_order}


class Target:
    """orn_target_create: framebuffer, accumulation and RNG state."""

    def __init__(self, scene, width, height, seed=0):
        h = C.c_void_p()
        check(lib().orn_target_create(scene.handle, int(width), int(height), int(seed),
                                      C.byref(h)), "orn_target_create")
        self._h = h
        self.scene = scene
        self.width, self.height = int(width), int(height)

    @property
    def handle(self):
        return self._h

    def render(self, params, iterations, wait=True):
        fn = lib().orn_render if wait else lib().orn_render_async
        check(fn(self._h, C.byref(params), int(iterations)), "orn_render")

    def wait(self):
        check(lib().orn_target_wait(self._h), "orn_target_wait")

    def set_tuning(self, regroup_lanes=48, workgroups_per_cu=0, bvh_width=2, leaf_regroup_lanes=16,
                   samples_per_job=16, split_percent=0, node_steps=4, workgroup_pool=0, cost_order=0):
        t = _tuning_struct(regroup_lanes, workgroups_per_cu, bvh_width, leaf_regroup_lanes,
                           samples_per_job, split_percent, node_steps, workgroup_pool, cost_order)
        check(lib().orn_target_set_tuning(self._h, C.byref(t)), "orn_target_set_tuning")

    def tuning(self):
        t = N.Tuning()
        check(lib().orn_target_get_tuning(self._h, C.byref(t)), "orn_target_get_tuning")
        return _tuning_dict(t)

    def autotune(self, params, iterations, allow_bvh4=False, cache_path=None, ignore_cache=False):
        """orn_target_autotune: times candidate tunings on renders of
        `iterations` spp, keeps the fastest; returns (tuning dict, ms)."""
        t, ms = N.Tuning(), C.c_float()
        flags = (N.TUNE_ALLOW_BVH4 if allow_bvh4 else 0) | (N.TUNE_IGNORE_CACHE if ignore_cache else 0)
        path = cache_path.encode() if cache_path else None
        check(lib().orn_target_autotune(self._h, C.byref(params), int(iterations), flags, path,
                                        C.byref(t), C.byref(ms)), "orn_target_autotune")
        return _tuning_dict(t), ms.value

    def last_render_ms(self):
        v = C.c_float()
        check(lib().orn_target_last_render_ms(self._h, C.byref(v)), "orn_target_last_render_ms")
        return v.value

    def _read(self, fn, dtype, comps, name):
        n = self.width * self.height
        out = np.empty((n, comps) if comps > 1 else n, dtype=dtype)
        check(fn(self._h, _ptr(out), n), name)
        return out

    def framebuffer(self):
        return self._read(lib().orn_get_framebuffer, np.float32, 4, "orn_get_framebuffer")

    def accumulation(self):
        return self._read(lib().orn_get_accumulation, np.float32, 4, "orn_get_accumulation")

    def rng_state(self):
        return self._read(lib().orn_get_rng_state, np.uint32, 1, "orn_get_rng_state")

    def pixel_costs(self):
        """Per-pixel GPU time (100 MHz ticks) recorded by the latest render
        with cost_order 1, row-major (orn_get_pixel_costs)."""
        return self._read(lib().orn_get_pixel_costs, np.uint32, 1, "orn_get_pixel_costs")

    def set_state(self, accumulation=None, rng=None):
        n = self.width * self.height
        a = None if accumulation is None else np.ascontiguousarray(accumulation, np.float32)
        r = None if rng is None else np.ascontiguousarray(rng, np.uint32)
        check(lib().orn_set_state(self._h, _ptr(a), _ptr(r), n), "orn_set_state")

    def destroy(self):
        if self._h:
            lib().orn_target_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Batch:
    """orn_batch_create: many images of many scenes rendered by one launch.

    `items` is a list of (Bvh arrays, width, height, seed)."""

    def __init__(self, device, items):
        keep = []
        arr = (N.BatchItem * len(items))()
        descs = []
        for i, (arrays, w, h, seed) in enumerate(items):
            d, k = scene_desc_arrays(arrays)
            keep.append(k)
            descs.append(d)
            arr[i].scene = C.addressof(d)
            arr[i].width, arr[i].height, arr[i].seed = int(w), int(h), int(seed)
        hnd = C.c_void_p()
        check(lib().orn_batch_create(device.handle, arr, len(items), C.byref(hnd)),
              "orn_batch_create")
        self._h = hnd
        self.device = device
        self.sizes = [(int(w), int(h)) for _, w, h, _ in items]

    @property
    def handle(self):
        return self._h

    def __len__(self):
        return len(self.sizes)

    def _args(self, params, iterations):
        if len(params) != len(self) or len(iterations) != len(self):
            raise ValueError("one params and one iteration count per image")
        p = (N.ConstantParams * len(self))(*params)
        it = (C.c_uint32 * len(self))(*[int(i) for i in iterations])
        return p, it

    def render(self, params, iterations, wait=True):
        p, it = self._args(params, iterations)
        fn = lib().orn_batch_render if wait else lib().orn_batch_render_async
        check(fn(self._h, p, it), "orn_batch_render")

    def wait(self):
        check(lib().orn_batch_wait(self._h), "orn_batch_wait")

    def set_tuning(self, **kw):
        t = _tuning_struct(**kw)
        check(lib().orn_batch_set_tuning(self._h, C.byref(t)), "orn_batch_set_tuning")

    def tuning(self):
        t = N.Tuning()
        check(lib().orn_batch_get_tuning(self._h, C.byref(t)), "orn_batch_get_tuning")
        return _tuning_dict(t)

    def autotune(self, params, iterations, allow_bvh4=False, cache_path=None, ignore_cache=False):
        p, it = self._args(params, iterations)
        t, ms = N.Tuning(), C.c_float()
        flags = (N.TUNE_ALLOW_BVH4 if allow_bvh4 else 0) | (N.TUNE_IGNORE_CACHE if ignore_cache else 0)
        path = cache_path.encode() if cache_path else None
        check(lib().orn_batch_autotune(self._h, p, it, flags, path, C.byref(t), C.byref(ms)),
              "orn_batch_autotune")
        return _tuning_dict(t), ms.value

    def last_render_ms(self):
        v = C.c_float()
        check(lib().orn_batch_last_render_ms(self._h, C.byref(v)), "orn_batch_last_render_ms")
        return v.value

    def _read(self, fn, item, dtype, comps, name):
        w, h = self.sizes[item]
        out = np.empty((w * h, comps) if comps > 1 else w * h, dtype=dtype)
        check(fn(self._h, int(item), _ptr(out), w * h), name)
        return out

    def framebuffer(self, item):
        return self._read(lib().orn_batch_get_framebuffer, item, np.float32, 4,
                          "orn_batch_get_framebuffer")

    def accumulation(self, item):
        return self._read(lib().orn_batch_get_accumulation, item, np.float32, 4,
                          "orn_batch_get_accumulation")

    def rng_state(self, item):
        return self._read(lib().orn_batch_get_rng_state, item, np.uint32, 1,
                          "orn_batch_get_rng_state")

    def set_state(self, item, accumulation=None, rng=None):
        w, h = self.sizes[item]
        a = None if accumulation is None else np.ascontiguousarray(accumulation, np.float32)
        r = None if rng is None else np.ascontiguousarray(rng, np.uint32)
        check(lib().orn_batch_set_state(self._h, int(item), _ptr(a), _ptr(r), w * h),
              "orn_batch_set_state")

    def destroy(self):
        if self._h:
            lib().orn_batch_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class PathTracer:
    """The reference's backend interface (wgpu_backend/path_tracer.zig:14-266).

    PathTracer(scene) takes ownership of the scene, as PathTracer.init does.
    """

    def __init__(self, scene, device=-1):
        h = C.c_void_p()
        check(lib().ornament_path_tracer_init(int(device), scene.handle, C.byref(h)),
              "PathTracer.init")
        scene._owned = False  # freed by deinit, like scene.deinit() in path_tracer.zig:127
        self._h = h
        self.scene = scene

    def deinit(self):
        if self._h:
            check(lib().ornament_path_tracer_deinit(self._h), "PathTracer.deinit")
            self.scene._h = None
        self._h = None

    def __del__(self):
        try:
            self.deinit()
        except Exception:
            pass

    def set_resolution(self, width, height):
        check(lib().ornament_path_tracer_set_resolution(self._h, int(width), int(height)),
              "setResolution")

    def render(self):
        check(lib().ornament_path_tracer_render(self._h), "render")

    def get_frame_buffer(self):
        w, h = self.resolution
        out = np.empty((w * h, 4), dtype=np.float32)
        check(lib().ornament_path_tracer_get_framebuffer(self._h, _ptr(out), w * h),
              "getFrameBuffer")
        return out

    def get_accumulation(self):
        w, h = self.resolution
        out = np.empty((w * h, 4), dtype=np.float32)
        check(lib().ornament_path_tracer_get_accumulation(self._h, _ptr(out), w * h),
              "getAccumulation")
        return out

    def last_render_ms(self):
        v = C.c_float()
        check(lib().ornament_path_tracer_last_render_ms(self._h, C.byref(v)), "last_render_ms")
        return v.value

    def constant_params(self):
        p = N.ConstantParams()
        check(lib().ornament_path_tracer_constant_params(self._h, C.byref(p)), "ConstantParams")
        return p

    # State (state.zig) -----------------------------------------------------
    def _get(self, fn, ctype):
        v = ctype()
        check(fn(self._h, C.byref(v)), fn.__name__)
        return v.value

    @property
    def depth(self):
        return self._get(lib().ornament_state_get_depth, C.c_uint32)

    @depth.setter
    def depth(self, v):
        check(lib().ornament_state_set_depth(self._h, int(v)), "setDepth")

    @property
    def iterations(self):
        return self._get(lib().ornament_state_get_iterations, C.c_uint32)

    @iterations.setter
    def iterations(self, v):
        check(lib().ornament_state_set_iterations(self._h, int(v)), "setIterations")

    @property
    def gamma(self):
        return self._get(lib().ornament_state_get_gamma, C.c_float)

    @gamma.setter
    def gamma(self, v):
        check(lib().ornament_state_set_gamma(self._h, float(v)), "setGamma")

    @property
    def flip_y(self):
        return bool(self._get(lib().ornament_state_get_flip_y, C.c_int))

    @flip_y.setter
    def flip_y(self, v):
        check(lib().ornament_state_set_flip_y(self._h, 1 if v else 0), "setFlipY")

    @property
    def ray_cast_epsilon(self):
        return self._get(lib().ornament_state_get_ray_cast_epsilon, C.c_float)

    @ray_cast_epsilon.setter
    def ray_cast_epsilon(self, v):
        check(lib().ornament_state_set_ray_cast_epsilon(self._h, float(v)), "setRayCastEpsilon")

    @property
    def current_iteration(self):
        return self._get(lib().ornament_state_get_current_iteration, C.c_float)

    @property
    def resolution(self):
        w, h = C.c_uint32(), C.c_uint32()
        check(lib().ornament_state_get_resolution(self._h, C.byref(w), C.byref(h)),
              "getResolution")
        return w.value, h.value
