"""GPU: orn_render_noalloc, the render entry point that allocates nothing,
against orn_render on a target (the existing entry point), through the C-ABI.

The image state lives in device buffers the caller allocated (here with
hipMalloc through ctypes); the launch's queue counters and stage words live
in the library's per-device scratch array. Two calls in a row, the second
with a smaller image, each bit for bit equal to a target's render with the
same tuning (split pixels on, so the stage words in the scratch carry the
hand-offs); an image larger than the scratch is refused before anything is
enqueued."""
import ctypes as C

import numpy as np
import pytest

import zig_ornament_amd as zo
from helpers import same_bits, zo_params
from zig_ornament_amd import _native as N
from zig_ornament_amd import workloads as W

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


class DeviceArray:
    """A device buffer holding a copy of a numpy array."""

    def __init__(self, h, host):
        self.h, self.shape, self.dtype = h, host.shape, host.dtype
        self.nbytes = host.nbytes
        self.ptr = C.c_void_p()
        assert h.hipMalloc(C.byref(self.ptr), self.nbytes) == 0
        assert h.hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, H2D) == 0

    def get(self):
        out = np.empty(self.shape, self.dtype)
        assert self.h.hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H) == 0
        return out

    def free(self):
        self.h.hipFree(self.ptr)


TUNING = dict(regroup_lanes=24, leaf_regroup_lanes=4, samples_per_job=2, split_percent=100, node_steps=3)


def noalloc_render(h, ds, spec, seed):
    n = spec.width * spec.height
    acc = DeviceArray(h, np.zeros((n, 4), np.float32))
    rng = DeviceArray(h, (np.arange(n, dtype=np.uint32) + np.uint32(seed * n)).astype(np.uint32))
    fb = DeviceArray(h, np.full((n, 4), -1.0, np.float32))
    t = zo._tuning_struct(**TUNING)
    p = zo_params(spec)
    st = N.lib().orn_render_noalloc(ds.handle, C.byref(p), spec.spp, C.byref(t), acc.ptr, rng.ptr, fb.ptr, None)
    assert st == 0, N.lib().orn_status_string(st)
    assert h.hipDeviceSynchronize() == 0
    out = acc.get(), rng.get(), fb.get()
    for b in (acc, rng, fb):
        b.free()
    return out


def target_render(ds, spec, seed):
    t = zo.Target(ds, spec.width, spec.height, seed)
    t.set_tuning(**TUNING)
    t.render(zo_params(spec), spec.spp)
    out = t.accumulation(), t.rng_state(), t.framebuffer()
    t.destroy()
    return out


def test_noalloc_render_matches_target_render(device):
    h = hip()
    big = W.cornell(width=96, height=80, spp=6)
    small = W.cornell(width=40, height=30, spp=3)
    small.flip_y, small.gamma = True, 2.2
    ds = zo.DeviceScene(device, zo.Bvh(W.build_scene(big)))
    for spec, seed in [(big, 0), (small, 2)]:  # the second call: a smaller image on the same scratch
        print(f"[noalloc] {spec.width}x{spec.height} {spec.spp} spp seed {seed}", flush=True)
        a_acc, a_rng, a_fb = noalloc_render(h, ds, spec, seed)
        b_acc, b_rng, b_fb = target_render(ds, spec, seed)
        assert np.all(a_acc[:, 3] == spec.spp)
        assert same_bits(a_acc, b_acc)
        assert np.array_equal(a_rng, b_rng)
        assert same_bits(a_fb, b_fb)


def test_noalloc_render_refuses_images_beyond_its_scratch(device):
    spec = W.cornell(width=2049, height=1024, spp=1)  # 2,098,176 pixels > 2^21
    ds = zo.DeviceScene(device, zo.Bvh(W.build_scene(W.cornell(width=8, height=8, spp=1))))
    p = zo_params(spec)
    st = N.lib().orn_render_noalloc(ds.handle, C.byref(p), 1, None, C.c_void_p(16), C.c_void_p(16),
                                    C.c_void_p(16), None)
    assert st == 9010  # ORN_ERROR_SCRATCH_TOO_SMALL: nothing was enqueued
    assert N.lib().orn_status_string(st) == b"ScratchTooSmall"
