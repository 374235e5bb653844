"""Scene files and image output (no GPU needed).

Scene files: a scene saved by the C++ host layer and loaded back builds the
same BVH arrays bit for bit; the independent Python reader of the same file
feeds the oracle's restatement of the scene factories and BVH build, which
must produce those arrays too (so the two readers pin each other and the
file carries everything the render depends on). Corrupt files are rejected
with a status, never read out of bounds.

Images: PFM is pinned by its published layout ("PF", width height, negative
scale = little-endian, RGB rows bottom to top) with a hand-assembled file;
OpenEXR by a reader written from the OpenEXR file-layout document (magic
20000630, version 2, attribute list, line offset table, uncompressed
scanline chunks with channels in alphabetical order). No OpenEXR library is
available here, so agreement with other EXR readers stays unpinned beyond
that layout."""
import os
import struct

import numpy as np
import pytest

import oracle_lib as O
import zig_ornament_amd as zo
from zig_ornament_amd import workloads as W

FIELDS = ["tlas_nodes", "blas_nodes", "normals", "normal_indices", "uvs", "uv_indices",
          "transforms", "materials"]


def small_specs():
    c3 = W.instanced_icospheres(width=32, height=18, spp=2, levels=2, instances=12,
                                texture_size=16, box=10.0)
    c2 = W.random_spheres(width=24, height=16, spp=2, grid=range(-3, 3))
    c2.flip_y, c2.gamma, c2.seed = True, 2.2, 7
    return {"C1": W.cornell(width=16, height=16, spp=2), "C2s": c2, "C3s": c3}


def same_arrays(a, b):
    return all(getattr(a, f).tobytes() == getattr(b, f).tobytes() for f in FIELDS)


@pytest.mark.parametrize("name", ["C1", "C2s", "C3s"])
def test_scene_file_round_trip(tmp_path, name):
    spec = small_specs()[name]
    path = str(tmp_path / f"{name}.orns")
    W.save_spec(spec, path)
    direct = zo.Bvh(W.build_scene(spec))
    scene, settings = zo.Scene.load(path)
    assert settings == {k: (np.float32(v).item() if isinstance(v, float) else v)
                        for k, v in W.spec_settings(spec).items()}
    loaded = zo.Bvh(scene)
    assert same_arrays(direct, loaded)
    assert [t["data"].tobytes() for t in direct.textures] == \
        [t["data"].tobytes() for t in loaded.textures]
    # the camera survives: same device camera bits
    assert bytes(scene.camera()) == bytes(W.build_scene_camera(spec))
    # the Python reader + the oracle's scene factories and BVH build agree
    spec2, settings2 = W.load_spec(path)
    assert settings2 == settings
    assert same_arrays(direct, O.build_bvh(spec2))
    # saving the loaded scene again gives the same bytes
    path2 = str(tmp_path / f"{name}_again.orns")
    scene.save(path2, settings)
    assert open(path, "rb").read() == open(path2, "rb").read()


def test_scene_file_without_settings(tmp_path):
    spec = small_specs()["C1"]
    path = str(tmp_path / "c1.orns")
    W.build_scene(spec).save(path)
    scene, settings = zo.Scene.load(path)
    assert settings is None
    assert same_arrays(zo.Bvh(W.build_scene(spec)), zo.Bvh(scene))


def test_corrupt_scene_files_are_rejected(tmp_path):
    spec = small_specs()["C3s"]
    path = str(tmp_path / "c3.orns")
    W.save_spec(spec, path)
    good = open(path, "rb").read()

    def status_of(data):
        p = str(tmp_path / "bad.orns")
        open(p, "wb").write(data)
        with pytest.raises(zo.OrnamentError) as e:
            zo.Scene.load(p)
        return e.value.name

    assert status_of(good[:-1]) == "InvalidFile"          # truncated
    assert status_of(good + b"\0") == "InvalidFile"       # trailing bytes
    assert status_of(b"ORNSCENF" + good[8:]) == "InvalidFile"
    assert status_of(good[:8] + struct.pack("<I", 2) + good[12:]) == "InvalidFile"  # version
    for cut in range(16, len(good), max(1, len(good) // 97)):
        assert status_of(good[:cut]) == "InvalidFile"
    # a material index out of range in the first instance record
    spec2, _ = W.load_spec(path)
    inst_at = len(good) - len(spec2.instances) * (4 + 64 + 8)
    bad = bytearray(good)
    struct.pack_into("<I", bad, inst_at + 4 + 64, 999)
    assert status_of(bytes(bad)) == "InvalidFile"
    struct.pack_into("<I", bad, inst_at + 4 + 64, 0)
    struct.pack_into("<I", bad, inst_at, 77)  # mesh index out of range
    assert status_of(bytes(bad)) == "InvalidFile"
    with pytest.raises(zo.OrnamentError) as e:
        zo.Scene.load(str(tmp_path / "missing.orns"))
    assert e.value.name == "IoError"


def _texture_only_scene(w, h, comps, bpc, data):
    """A scene file with no camera settings and one texture record."""
    head = b"ORNSCENE" + struct.pack("<II", 1, 0) + struct.pack("<13f", *([0.0] * 13))
    head += struct.pack("<5I", 1, 0, 0, 0, 0)
    rec = struct.pack("<5IfQ", w, h, comps, bpc, 1 if bpc == 2 else 0, 1.0, len(data)) + data
    return head + rec + b"\0" * (-len(rec) % 4)


def test_texture_sizes_that_overflow_are_rejected(tmp_path):
    """4294705160 x 536903681 RGBA16F is 64 bytes modulo 2^64: without a bound
    on the dimensions it would pass the size check and the kernel would index
    texels 2^61 bytes past a 64-byte pool."""
    assert (4294705160 * 536903681 * 4 * 2) % 2**64 == 64
    p = str(tmp_path / "overflow.orns")
    open(p, "wb").write(_texture_only_scene(4294705160, 536903681, 4, 2, bytes(64)))
    with pytest.raises(zo.OrnamentError) as e:
        zo.Scene.load(p)
    assert e.value.name == "InvalidFile"
    # the bound itself: WebGPU's default maxTextureDimension2D, 8192
    s = zo.Scene()
    with pytest.raises(zo.OrnamentError) as e:
        s.create_texture(np.zeros(8193 * 4, np.uint8), 8193, 1, 4)
    assert e.value.name == "UnsupportedTexture"
    s.create_texture(np.zeros(8192 * 4, np.uint8), 8192, 1, 4)
    # a well-formed texture-only file loads
    open(p, "wb").write(_texture_only_scene(2, 2, 4, 1, bytes(16)))
    zo.Scene.load(p)


def test_pfm_layout_and_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    w, h = 5, 3
    rgba = rng.random((w * h, 4), dtype=np.float32)
    path = str(tmp_path / "a.pfm")
    zo.write_pfm(path, rgba, w, h)
    data = open(path, "rb").read()
    head = b"PF\n5 3\n-1.0\n"
    assert data[:len(head)] == head
    body = np.frombuffer(data[len(head):], "<f4").reshape(h * w, 3)
    assert np.array_equal(body, rgba[:, :3])  # row 0 (bottom) first, as PFM stores rows
    back, bw, bh = zo.read_pfm(path)
    assert (bw, bh) == (w, h)
    assert np.array_equal(back[:, :3], rgba[:, :3]) and np.all(back[:, 3] == 1.0)


def test_pfm_reader_accepts_big_endian_and_grey(tmp_path):
    vals = np.array([0.5, -2.0, 3.25, 1e-3, 7.0, 8.5], np.float32)
    p = str(tmp_path / "be.pfm")
    open(p, "wb").write(b"PF\n2 1\n1.0\n" + vals.astype(">f4").tobytes())
    img, w, h = zo.read_pfm(p)
    assert (w, h) == (2, 1)
    assert np.array_equal(img[:, :3].reshape(-1), vals)
    p = str(tmp_path / "grey.pfm")
    open(p, "wb").write(b"Pf 3 1 -1\n" + vals[:3].astype("<f4").tobytes())
    img, w, h = zo.read_pfm(p)
    assert (w, h) == (3, 1)
    assert np.array_equal(img[:, 0], vals[:3]) and np.array_equal(img[:, 2], vals[:3])
    p = str(tmp_path / "short.pfm")
    open(p, "wb").write(b"PF\n2 2\n-1.0\n" + b"\0" * 40)
    with pytest.raises(zo.OrnamentError) as e:
        zo.read_pfm(p)
    assert e.value.name == "InvalidFile"


def read_exr(path):
    """Minimal OpenEXR reader for single-part, uncompressed scanline files
    with FLOAT channels, written from the file-layout document."""
    b = open(path, "rb").read()
    magic, version = struct.unpack_from("<ii", b, 0)
    assert magic == 20000630 and version & 0xff == 2 and not version & 0x200  # not tiled
    at = 8
    attrs = {}
    while b[at] != 0:
        name_end = b.index(b"\0", at)
        name = b[at:name_end].decode()
        type_end = b.index(b"\0", name_end + 1)
        typ = b[name_end + 1:type_end].decode()
        size, = struct.unpack_from("<i", b, type_end + 1)
        attrs[name] = (typ, b[type_end + 5:type_end + 5 + size])
        at = type_end + 5 + size
    at += 1
    chans = []
    raw = attrs["channels"][1]
    i = 0
    while raw[i] != 0:
        e = raw.index(b"\0", i)
        name = raw[i:e].decode()
        ptype, _, xs, ys = struct.unpack_from("<iIii", raw, e + 1)
        assert ptype == 2 and xs == 1 and ys == 1
        chans.append(name)
        i = e + 17
    assert chans == sorted(chans)
    assert attrs["compression"][1] == b"\0"
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    assert attrs["displayWindow"][1] == attrs["dataWindow"][1]
    w, h = x1 - x0 + 1, y1 - y0 + 1
    offsets = struct.unpack_from(f"<{h}Q", b, at)
    img = {c: np.zeros((h, w), np.float32) for c in chans}
    for k, off in enumerate(offsets):
        y, n = struct.unpack_from("<iI", b, off)
        assert n == 4 * w * len(chans)
        row = np.frombuffer(b, "<f4", w * len(chans), off + 8).reshape(len(chans), w)
        for ci, c in enumerate(chans):
            img[c][y - y0] = row[ci]
    return img, w, h, attrs


def test_exr_layout_and_values(tmp_path):
    rng = np.random.default_rng(2)
    w, h = 7, 4
    rgba = rng.random((w * h, 4), dtype=np.float32) * 10
    path = str(tmp_path / "a.exr")
    zo.write_exr(path, rgba, w, h)
    img, ew, eh, attrs = read_exr(path)
    assert (ew, eh) == (w, h)
    for a in ("channels", "compression", "dataWindow", "displayWindow", "lineOrder",
              "pixelAspectRatio", "screenWindowCenter", "screenWindowWidth"):
        assert a in attrs
    assert attrs["lineOrder"][1] == b"\0"
    grid = rgba.reshape(h, w, 4)[::-1]  # EXR row 0 is the top of the picture
    for ci, c in enumerate("RGBA"):
        assert np.array_equal(img[c], grid[:, :, ci]), c


def test_image_writers_reject_bad_arguments(tmp_path):
    with pytest.raises(ValueError):
        zo.write_pfm(str(tmp_path / "x.pfm"), np.zeros((3, 4), np.float32), 2, 2)
    rc = zo.lib().ornament_image_write_exr(os.fsencode(str(tmp_path / "x.exr")), None, 2, 2)
    assert rc == 1
    rc = zo.lib().ornament_image_write_pfm(os.fsencode(str(tmp_path / "nodir" / "x.pfm")),
                                           np.zeros(16, np.float32).ctypes.data, 2, 2)
    assert rc == 9008


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_golden_reference_image_is_the_oracles_render():
    """The committed 4,096-spp image of the committed scene file is what the
    oracle renders from that file (both made by make_reference_images.py)."""
    import sys
    sys.path.insert(0, GOLDEN)
    from make_reference_images import IMAGE, SCENE, oracle_image
    fb, spec = oracle_image(SCENE)
    ref, w, h = zo.read_pfm(IMAGE)
    assert (w, h) == (spec.width, spec.height) == (64, 64) and spec.spp == 4096
    assert np.array_equal(ref[:, :3], fb[:, :3])
