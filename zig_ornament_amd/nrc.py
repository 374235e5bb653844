"""Neural radiance cache (orn_nrc_*): a hash-grid encoded 5 x 64 MLP on BF16
MFMA that terminates paths early, trained online on the tracer's own
radiance samples. See include/ornament_hip.h for the method."""
import ctypes as C

import numpy as np

from . import _native as N
from ._native import check, lib

RECORD_DTYPE = np.dtype([("pos", "<f4", 3), ("pixel", "<u4"), ("dir", "<f4", 3), ("roughness", "<f4"),
                         ("normal", "<f4", 3), ("type", "<u4"), ("albedo", "<f4", 3), ("flags", "<u4"),
                         ("weight", "<f4", 3), ("pad", "<u4")])
assert RECORD_DTYPE.itemsize == 80

MLP_PARAMS = 5 * 64 * 64 + 64 * 16


def default_config(**overrides):
    c = N.NrcConfig()
    check(lib().orn_nrc_default_config(C.byref(c)), "orn_nrc_default_config")
    for k, v in overrides.items():
        setattr(c, k, v)
    return c


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class Nrc:
    def __init__(self, device_scene, **config):
        h = C.c_void_p()
        cfg = default_config(**config)
        check(lib().orn_nrc_create(device_scene.handle, C.byref(cfg), C.byref(h)), "orn_nrc_create")
        self._h = h
        self.scene = device_scene
        self.config = {f: getattr(cfg, f) for f, _ in N.NrcConfig._fields_ if f != "reserved"}

    @property
    def handle(self):
        return self._h

    def info(self):
        i = N.NrcInfo()
        check(lib().orn_nrc_get_info(self._h, C.byref(i)), "orn_nrc_get_info")
        return dict(levels=i.levels, log2_table=i.log2_table, resolution=list(i.resolution),
                    dense=list(i.dense), bounds_lo=list(i.bounds_lo), inv_extent=list(i.inv_extent),
                    mlp_params=i.mlp_params, grid_params=i.grid_params, adam_steps=i.adam_steps)

    def get_params(self):
        gp = self.info()["grid_params"]
        mlp = np.empty(MLP_PARAMS, np.float32)
        grid = np.empty(gp, np.float32)
        check(lib().orn_nrc_get_params(self._h, _p(mlp), mlp.size, _p(grid), grid.size),
              "orn_nrc_get_params")
        return mlp, grid

    def set_params(self, mlp=None, grid=None):
        m = None if mlp is None else np.ascontiguousarray(mlp, np.float32)
        g = None if grid is None else np.ascontiguousarray(grid, np.float32)
        check(lib().orn_nrc_set_params(self._h, _p(m), 0 if m is None else m.size, _p(g),
                                       0 if g is None else g.size), "orn_nrc_set_params")

    def query(self, records, with_inputs=False):
        r = np.ascontiguousarray(records, RECORD_DTYPE)
        rgb = np.empty((len(r), 3), np.float32)
        inputs = np.empty((len(r), 64), np.float32) if with_inputs else None
        check(lib().orn_nrc_query(self._h, _p(r), len(r), _p(rgb), _p(inputs)), "orn_nrc_query")
        return (rgb, inputs) if with_inputs else rgb

    def gradients(self, records, with_grid=True):
        r = np.ascontiguousarray(records, RECORD_DTYPE)
        loss = C.c_float()
        mg = np.empty(MLP_PARAMS, np.float32)
        gg = np.empty(self.info()["grid_params"], np.float32) if with_grid else None
        check(lib().orn_nrc_gradients(self._h, _p(r), len(r), C.byref(loss), _p(mg), _p(gg),
                                      0 if gg is None else gg.size), "orn_nrc_gradients")
        return loss.value, mg, gg

    def step(self, records):
        r = np.ascontiguousarray(records, RECORD_DTYPE)
        loss = C.c_float()
        check(lib().orn_nrc_step(self._h, _p(r), len(r), C.byref(loss)), "orn_nrc_step")
        return loss.value

    def render(self, target, params, iterations):
        check(lib().orn_nrc_render(self._h, target.handle, C.byref(params), int(iterations)),
              "orn_nrc_render")

    def train(self, params, frame, steps=4, want_loss=True):
        loss = C.c_float()
        check(lib().orn_nrc_train(self._h, C.byref(params), int(frame), int(steps),
                                  C.byref(loss) if want_loss else None), "orn_nrc_train")
        return loss.value if want_loss else None

    def frame(self, target, params, iterations, frame, steps=1, want_loss=False):
        """orn_nrc_frame: training paths beside the render, then Adam steps;
        returns (frame GPU ms, loss or None)."""
        loss = C.c_float()
        check(lib().orn_nrc_frame(self._h, target.handle, C.byref(params), int(iterations), int(frame),
                                  int(steps), C.byref(loss) if want_loss else None), "orn_nrc_frame")
        ms = C.c_float()
        check(lib().orn_nrc_frame_ms(self._h, C.byref(ms)), "orn_nrc_frame_ms")
        return ms.value, (loss.value if want_loss else None)

    def last_ms(self):
        r, t = C.c_float(), C.c_float()
        check(lib().orn_nrc_last_ms(self._h, C.byref(r), C.byref(t)), "orn_nrc_last_ms")
        return r.value, t.value

    def destroy(self):
        if self._h:
            lib().orn_nrc_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
