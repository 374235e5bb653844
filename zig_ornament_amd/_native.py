"""ctypes declarations of include/ornament_hip.h and include/ornament.h.

The shared library is built in-tree by zig_ornament_amd/build.py. Loading
fails loudly when it is missing: there is no fallback path.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# ORN_LIB selects another build of the library (the sanitizer build of
# zig_ornament_amd/build.py --sanitize)
LIB_PATH = os.environ.get("ORN_LIB") or os.path.join(HERE, "libornament_hip.so")

u32 = C.c_uint32
f32 = C.c_float
P = C.POINTER


class BvhNode(C.Structure):  # gpu_structs.zig:22-31
    _fields_ = [("lmin_or_v0", f32 * 3), ("left_or_custom_id", u32),
                ("lmax_or_v1", f32 * 3), ("right_or_material_index", u32),
                ("rmin_or_v2", f32 * 3), ("node_type", u32),
                ("rmax_or_v3", f32 * 3), ("transform_id", u32)]


class Material(C.Structure):  # gpu_structs.zig:33-59
    _fields_ = [("albedo", f32 * 3), ("albedo_texture_index", u32), ("fuzz", f32),
                ("ior", f32), ("type", u32), ("padding", u32)]


class Camera(C.Structure):  # gpu_structs.zig:88-117
    _fields_ = [("origin", f32 * 3), ("lens_radius", f32),
                ("lower_left_corner", f32 * 3), ("padding0", u32),
                ("horizontal", f32 * 3), ("padding1", u32),
                ("vertical", f32 * 3), ("padding2", u32),
                ("u", f32 * 3), ("padding3", u32),
                ("v", f32 * 3), ("padding4", u32),
                ("w", f32 * 3), ("padding5", u32)]


class ConstantParams(C.Structure):  # gpu_structs.zig:61-86
    _fields_ = [("camera", Camera), ("depth", u32), ("width", u32), ("height", u32),
                ("flip_y", u32), ("inverted_gamma", f32), ("ray_cast_epsilon", f32),
                ("textures_count", u32), ("current_iteration", f32)]


class TextureDesc(C.Structure):
    _fields_ = [("width", u32), ("height", u32), ("num_components", u32),
                ("bytes_per_component", u32), ("is_hdr", u32), ("padding", u32),
                ("data", C.c_void_p)]


class SceneDesc(C.Structure):
    _fields_ = [("tlas_nodes", C.c_void_p), ("tlas_nodes_len", C.c_size_t),
                ("blas_nodes", C.c_void_p), ("blas_nodes_len", C.c_size_t),
                ("normals", C.c_void_p), ("normals_len", C.c_size_t),
                ("normal_indices", C.c_void_p), ("normal_indices_len", C.c_size_t),
                ("uvs", C.c_void_p), ("uvs_len", C.c_size_t),
                ("uv_indices", C.c_void_p), ("uv_indices_len", C.c_size_t),
                ("transforms", C.c_void_p), ("transforms_len", C.c_size_t),
                ("materials", C.c_void_p), ("materials_len", C.c_size_t),
                ("textures", C.c_void_p), ("textures_len", C.c_size_t)]


class Tuning(C.Structure):
    _fields_ = [("regroup_lanes", u32), ("workgroups_per_cu", u32), ("bvh_width", u32),
                ("leaf_regroup_lanes", u32), ("samples_per_job", u32), ("split_percent", u32), ("node_steps", u32),
                ("workgroup_pool", u32), ("cost_order", u32)]


class RenderSettings(C.Structure):
    _fields_ = [("width", u32), ("height", u32), ("iterations", u32), ("depth", u32),
                ("gamma", f32), ("ray_cast_epsilon", f32), ("flip_y", u32), ("seed", u32)]


class NrcConfig(C.Structure):
    _fields_ = [("log2_table", u32), ("base_resolution", u32), ("max_resolution", u32),
                ("terminate_bounce", u32), ("train_paths", u32), ("train_batch", u32),
                ("learning_rate", f32), ("beta1", f32), ("beta2", f32), ("epsilon", f32),
                ("seed", u32), ("reserved", u32 * 5)]


class NrcInfo(C.Structure):
    _fields_ = [("levels", u32), ("log2_table", u32), ("resolution", u32 * 16), ("dense", u32 * 16),
                ("bounds_lo", f32 * 3), ("inv_extent", f32 * 3), ("mlp_params", u32),
                ("grid_params", C.c_uint64), ("adam_steps", u32), ("reserved", u32 * 7)]


class BatchItem(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("width", u32), ("height", u32), ("seed", u32),
                ("reserved", u32)]


TUNE_ALLOW_BVH4 = 1
TUNE_IGNORE_CACHE = 2


class Hit(C.Structure):
    _fields_ = [("hit", u32), ("t", f32), ("material_index", u32), ("node_type", u32),
                ("inverted_transform_id", u32), ("tri_id", u32), ("uv", f32 * 2)]


assert C.sizeof(BvhNode) == 64 and C.sizeof(Material) == 32
assert C.sizeof(Camera) == 112 and C.sizeof(ConstantParams) == 144
assert C.sizeof(Hit) == 32

vp = C.c_void_p
fp = P(f32)
up = P(u32)
st = C.c_int
sz = C.c_size_t

# (name, restype, argtypes) of every exported function
DEVICE_API = [
    ("orn_status_string", C.c_char_p, [st]),
    ("orn_device_count", st, [P(C.c_int)]),
    ("orn_device_open", st, [C.c_int, P(vp)]),
    ("orn_device_close", st, [vp]),
    ("orn_device_info", st, [vp, C.c_char_p, sz, P(C.c_int)]),
    ("orn_scene_upload", st, [vp, P(SceneDesc), P(vp)]),
    ("orn_scene_destroy", st, [vp]),
    ("orn_scene_validate", st, [P(SceneDesc), up]),
    ("orn_scene_stack_depth", st, [vp, up]),
    ("orn_scene_set_trace_width", st, [vp, u32]),
    ("orn_target_create", st, [vp, u32, u32, u32, P(vp)]),
    ("orn_target_destroy", st, [vp]),
    ("orn_render", st, [vp, P(ConstantParams), u32]),
    ("orn_render_async", st, [vp, P(ConstantParams), u32]),
    ("orn_target_wait", st, [vp]),
    ("orn_render_noalloc", st, [vp, P(ConstantParams), u32, P(Tuning), vp, vp, vp, vp]),
    ("orn_target_last_render_ms", st, [vp, fp]),
    ("orn_target_set_tuning", st, [vp, P(Tuning)]),
    ("orn_target_get_tuning", st, [vp, P(Tuning)]),
    ("orn_get_framebuffer", st, [vp, vp, sz]),
    ("orn_get_accumulation", st, [vp, vp, sz]),
    ("orn_get_rng_state", st, [vp, vp, sz]),
    ("orn_get_pixel_costs", st, [vp, vp, sz]),
    ("orn_set_state", st, [vp, vp, vp, sz]),
    ("orn_trace_rays", st, [vp, f32, vp, sz, vp]),
    ("orn_trace_rays_packet", st, [vp, f32, vp, sz, vp]),
    ("orn_eval_math", st, [vp, C.c_int, vp, vp, vp, sz]),
    ("orn_calibration_copy", st, [vp, sz, fp]),
    ("orn_batch_create", st, [vp, P(BatchItem), sz, P(vp)]),
    ("orn_batch_destroy", st, [vp]),
    ("orn_batch_render", st, [vp, P(ConstantParams), up]),
    ("orn_batch_render_async", st, [vp, P(ConstantParams), up]),
    ("orn_batch_wait", st, [vp]),
    ("orn_batch_set_tuning", st, [vp, P(Tuning)]),
    ("orn_batch_get_tuning", st, [vp, P(Tuning)]),
    ("orn_batch_last_render_ms", st, [vp, fp]),
    ("orn_batch_get_framebuffer", st, [vp, sz, vp, sz]),
    ("orn_batch_get_accumulation", st, [vp, sz, vp, sz]),
    ("orn_batch_get_rng_state", st, [vp, sz, vp, sz]),
    ("orn_batch_set_state", st, [vp, sz, vp, vp, sz]),
    ("orn_target_autotune", st, [vp, P(ConstantParams), u32, u32, C.c_char_p, P(Tuning), fp]),
    ("orn_batch_autotune", st, [vp, P(ConstantParams), up, u32, C.c_char_p, P(Tuning), fp]),
    ("orn_nrc_default_config", st, [P(NrcConfig)]),
    ("orn_nrc_create", st, [vp, P(NrcConfig), P(vp)]),
    ("orn_nrc_destroy", st, [vp]),
    ("orn_nrc_get_info", st, [vp, P(NrcInfo)]),
    ("orn_nrc_render", st, [vp, vp, P(ConstantParams), u32]),
    ("orn_nrc_train", st, [vp, P(ConstantParams), u32, u32, fp]),
    ("orn_nrc_last_ms", st, [vp, fp, fp]),
    ("orn_nrc_frame", st, [vp, vp, P(ConstantParams), u32, u32, u32, fp]),
    ("orn_nrc_frame_ms", st, [vp, fp]),
    ("orn_nrc_get_params", st, [vp, vp, sz, vp, sz]),
    ("orn_nrc_set_params", st, [vp, vp, sz, vp, sz]),
    ("orn_nrc_query", st, [vp, vp, sz, vp, vp]),
    ("orn_nrc_gradients", st, [vp, vp, sz, fp, vp, vp, sz]),
    ("orn_nrc_step", st, [vp, vp, sz, fp]),
]

HOST_API = [
    ("ornament_scene_init", st, [P(vp)]),
    ("ornament_scene_deinit", None, [vp]),
    ("ornament_scene_lambertian", st, [vp, fp, P(vp)]),
    ("ornament_scene_lambertian_texture", st, [vp, vp, P(vp)]),
    ("ornament_scene_metal", st, [vp, fp, f32, P(vp)]),
    ("ornament_scene_metal_texture", st, [vp, vp, f32, P(vp)]),
    ("ornament_scene_dielectric", st, [vp, f32, P(vp)]),
    ("ornament_scene_diffuse_light", st, [vp, fp, P(vp)]),
    ("ornament_scene_diffuse_light_texture", st, [vp, vp, P(vp)]),
    ("ornament_scene_create_sphere", st, [vp, fp, f32, vp, P(vp)]),
    ("ornament_scene_create_sphere_mesh", st, [vp, fp, f32, vp, P(vp)]),
    ("ornament_scene_create_plane_mesh", st, [vp, fp, f32, f32, fp, vp, P(vp)]),
    ("ornament_scene_create_mesh", st, [vp, vp, sz, vp, sz, vp, sz, vp, sz, vp, sz, vp, sz,
                                        fp, vp, P(vp)]),
    ("ornament_scene_create_mesh_instance", st, [vp, vp, fp, vp, P(vp)]),
    ("ornament_scene_create_texture", st, [vp, vp, sz, u32, u32, u32, u32, C.c_int, f32,
                                           P(vp)]),
    ("ornament_scene_attach_sphere", st, [vp, vp]),
    ("ornament_scene_attach_mesh", st, [vp, vp]),
    ("ornament_scene_attach_mesh_instance", st, [vp, vp]),
    ("ornament_scene_counts", st, [vp, up]),
    ("ornament_scene_set_camera", st, [vp, fp, fp, fp, f32, f32, f32, f32]),
    ("ornament_camera_set_look_at", st, [vp, fp, fp, fp]),
    ("ornament_camera_set_aspect_ratio", st, [vp, f32]),
    ("ornament_camera_get", st, [vp, P(Camera)]),
    ("ornament_camera_get_dirty", st, [vp, P(C.c_int)]),
    ("ornament_bvh_init", st, [vp, C.c_int, P(vp)]),
    ("ornament_bvh_shared_generator", st, [C.c_int]),
    ("ornament_bvh_deinit", None, [vp]),
    ("ornament_bvh_arrays", st, [vp, P(SceneDesc)]),
    ("ornament_bvh_texture_sources", st, [vp, up, sz]),
    ("ornament_path_tracer_init", st, [C.c_int, vp, P(vp)]),
    ("ornament_path_tracer_deinit", st, [vp]),
    ("ornament_path_tracer_scene", vp, [vp]),
    ("ornament_path_tracer_set_resolution", st, [vp, u32, u32]),
    ("ornament_path_tracer_render", st, [vp]),
    ("ornament_path_tracer_get_framebuffer", st, [vp, vp, sz]),
    ("ornament_path_tracer_get_accumulation", st, [vp, vp, sz]),
    ("ornament_path_tracer_last_render_ms", st, [vp, fp]),
    ("ornament_state_set_depth", st, [vp, u32]),
    ("ornament_state_get_depth", st, [vp, up]),
    ("ornament_state_set_iterations", st, [vp, u32]),
    ("ornament_state_get_iterations", st, [vp, up]),
    ("ornament_state_set_gamma", st, [vp, f32]),
    ("ornament_state_get_gamma", st, [vp, fp]),
    ("ornament_state_set_flip_y", st, [vp, C.c_int]),
    ("ornament_state_get_flip_y", st, [vp, P(C.c_int)]),
    ("ornament_state_set_ray_cast_epsilon", st, [vp, f32]),
    ("ornament_state_get_ray_cast_epsilon", st, [vp, fp]),
    ("ornament_state_get_resolution", st, [vp, up, up]),
    ("ornament_state_get_current_iteration", st, [vp, fp]),
    ("ornament_path_tracer_constant_params", st, [vp, P(ConstantParams)]),
    ("ornament_scene_save", st, [vp, P(RenderSettings), C.c_char_p]),
    ("ornament_scene_load", st, [C.c_char_p, P(vp), P(RenderSettings), P(C.c_int)]),
    ("ornament_image_write_pfm", st, [C.c_char_p, vp, u32, u32]),
    ("ornament_image_read_pfm", st, [C.c_char_p, vp, sz, up, up]),
    ("ornament_image_write_exr", st, [C.c_char_p, vp, u32, u32]),
]

ALL_API = DEVICE_API + HOST_API

_lib = None


class OrnamentError(RuntimeError):
    def __init__(self, status, where):
        self.status = status
        name = lib().orn_status_string(status)
        name = name.decode() if name else str(status)
        super().__init__(f"{where} failed: {name} ({status})")
        self.name = name


def lib():
    """The loaded libornament_hip.so (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with zig_ornament_amd/build.py "
                "(or __graft_entry__.build()) before use")
        L = C.CDLL(LIB_PATH)
        for name, res, args in ALL_API:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status, where):
    if status != 0:
        raise OrnamentError(status, where)
