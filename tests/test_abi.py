"""The C-ABI library: it loads without a GPU, exports every function the
headers declare, and its host-side error paths behave (no GPU needed)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import zig_ornament_amd as zo
from zig_ornament_amd import _native as N
from zig_ornament_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in ("ornament_hip.h", "ornament.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b((?:orn|ornament)_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_every_declared_symbol_is_exported():
    lib = zo.lib()
    names = declared_functions()
    assert len(names) > 60
    for n in names:
        assert hasattr(lib, n), n
    # and the Python declarations cover exactly those
    assert sorted(n for n, _, _ in N.ALL_API) == names


def test_status_strings():
    lib = zo.lib()
    assert lib.orn_status_string(0) == b"Success"
    assert lib.orn_status_string(9001) == b"EmptyScene"
    assert lib.orn_status_string(2) == b"hipErrorOutOfMemory"


def test_null_handles_are_rejected():
    lib = zo.lib()
    assert lib.orn_device_close(None) == 9006
    assert lib.orn_scene_destroy(None) == 9006
    assert lib.orn_target_destroy(None) == 9006
    assert lib.orn_render(None, None, 1) == 9006
    assert lib.orn_get_framebuffer(None, None, 0) == 9006
    assert lib.ornament_path_tracer_render(None) == 9006


def _desc(prod):
    d, keep = zo.scene_desc_arrays(prod)
    return d, keep


def test_validation_accepts_real_scenes_and_sizes_the_stack():
    for spec in (W.cornell(), W.random_spheres(width=8, height=8)):
        prod = zo.Bvh(W.build_scene(spec))
        d, keep = _desc(prod)
        slots = C.c_uint32()
        assert zo.lib().orn_scene_validate(C.byref(d), C.byref(slots)) == 0
        assert 4 <= slots.value <= 63


def test_validation_rejects_out_of_range_indices():
    prod = zo.Bvh(W.build_scene(W.cornell()))

    def status(mutate):
        import copy
        p = copy.deepcopy(prod)
        mutate(p)
        d, keep = _desc(p)
        return zo.lib().orn_scene_validate(C.byref(d), None)

    assert status(lambda p: None) == 0
    def bad_child(p):
        root = len(p.tlas_nodes) - 1
        p.tlas_nodes[root]["left_or_custom_id"] = root  # a cycle
    assert status(bad_child) == 9003
    def bad_transform(p):
        leaf = int(np.nonzero(p.tlas_nodes["node_type"] == 1)[0][0])
        p.tlas_nodes[leaf]["transform_id"] = 10 ** 6
    assert status(bad_transform) == 9003
    def bad_normal(p):
        p.normal_indices[0] = len(p.normals)
    assert status(bad_normal) == 9003
    def bad_blas_root(p):
        leaf = int(np.nonzero(p.tlas_nodes["node_type"] == 2)[0][0])
        p.tlas_nodes[leaf]["left_or_custom_id"] = len(p.blas_nodes)
    assert status(bad_blas_root) == 9003
    def bad_material(p):
        leaf = int(np.nonzero(p.tlas_nodes["node_type"] == 2)[0][0])
        p.tlas_nodes[leaf]["right_or_material_index"] = len(p.materials)
    assert status(bad_material) == 9003
    def empty(p):
        p.tlas_nodes = p.tlas_nodes[:0]
    assert status(empty) == 9001


def test_validation_rejects_unsupported_textures():
    spec = W.instanced_icospheres(levels=1, instances=2, texture_size=4)
    prod = zo.Bvh(W.build_scene(spec))
    prod.textures[0] = dict(prod.textures[0], num_components=3)
    d, keep = _desc(prod)
    assert zo.lib().orn_scene_validate(C.byref(d), None) == 9005


def test_host_errors_mirror_reference_panics():
    s = zo.Scene()
    with pytest.raises(zo.OrnamentError) as e:
        zo.Bvh(s)
    assert e.value.name == "EmptyScene"  # bvh.zig:35-37
    m = s.lambertian((0.5, 0.5, 0.5))
    s.create_mesh(np.zeros((3, 4), np.float32), [0, 1], np.zeros((3, 4), np.float32), [0, 1],
                  None, None, np.eye(4), m)
    with pytest.raises(zo.OrnamentError) as e:
        zo.Bvh(s)
    assert e.value.name == "EmptyBvh"  # bvh.zig:222-223


def test_no_device_here_is_reported(monkeypatch):
    if zo.Device.count() > 0:
        pytest.skip("a GPU is visible")
    lib = zo.lib()
    h = C.c_void_p()
    assert lib.orn_device_open(0, C.byref(h)) == 100
    s = zo.Scene()
    s.create_sphere((0, 0, 0), 1.0, s.lambertian((1, 1, 1)))
    assert lib.ornament_path_tracer_init(0, s.handle, C.byref(h)) == 100


def test_scene_api_counts_and_camera():
    s = zo.Scene()
    m = s.lambertian((0.5, 0.5, 0.5))
    s.create_sphere((0, 0, 0), 1.0, m)
    s.create_sphere_mesh((0, 0, 0), 1.0, m)
    mesh = s.create_plane_mesh((0, 0, 0), 1, 1, (0, 1, 0), m)
    s.create_mesh_instance(mesh, np.eye(4), m)
    s.create_texture(np.zeros(16, np.uint8), 2, 2, 4)
    assert s.counts() == dict(spheres=1, meshes=2, mesh_instances=1, materials=1, textures=1)
    assert s.camera_dirty()
    s.set_look_at((0, 0, 5), (0, 0, 0), (0, 1, 0))
    cam = s.camera()
    assert np.allclose(cam.w, [0, 0, 1]) and cam.origin[2] == 5.0
    # createSphereMesh: 60 x 30 UV sphere (scene.zig:140-233)
    b = zo.Bvh(s)
    assert len(b.normals) == 4 + (1 + 29 * 60 + 1)


def test_shipped_tuning_cache_entries_are_valid():
    """Every entry of the shipped cache names a scene class the library
    builds keys for and holds a tuning orn_target_set_tuning accepts (ranges
    of orn_tuning in include/ornament_hip.h); an invalid entry would be
    skipped and the class measured again."""
    import json
    cache = json.load(open(zo.TUNING_CACHE))
    assert len(cache) >= 4
    key = re.compile(r"^gfx[0-9a-f]+[^|]*\|features[0-7]\|tlas2\^\d+\|blas2\^\d+\|px2\^\d+\|spp2\^\d+\|depth\d+"
                     r"\|images2\^\d+\|bvh4[01]$")
    for k, e in cache.items():
        assert key.match(k), k
        assert 1 <= e["regroup_lanes"] <= 64 and 1 <= e["leaf_regroup_lanes"] <= 64, k
        assert e["workgroups_per_cu"] <= 64 and e["bvh_width"] in (2, 4), k
        assert e["samples_per_job"] <= 0xFFFF and e["split_percent"] <= 100, k
        assert 0 <= e.get("node_steps", 0) <= 16, k
        assert e.get("workgroup_pool", 0) in (0, 1), k
        assert e.get("cost_order", 0) in (0, 1), k
        assert e["ms"] > 0, k
