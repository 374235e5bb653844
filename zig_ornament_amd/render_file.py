"""Renders a scene file to a PFM or OpenEXR image on the GPU.

    python -m zig_ornament_amd.render_file SCENE.orns OUT.pfm [--spp N]
        [--width W --height H] [--device D]

The scene file's render settings (resolution, samples per pixel, depth,
gamma, flip_y, ray_cast_epsilon, RNG seed) apply unless overridden; one
render() of `spp` iterations runs from the seeded state and the
post-processed framebuffer is written (PFM: RGB rows bottom to top; EXR:
RGBA, rows top to bottom). This is how reference images (4,096 spp and up)
are made reproducibly: the same file renders to the same bits on any MI355X.
"""
import argparse
import sys

from . import Device, DeviceScene, Target, Scene, Bvh, write_exr, write_pfm
from . import _native as N


def render(path, spp=None, width=None, height=None, device=0):
    """(rgba, width, height, milliseconds) of one render of a scene file."""
    scene, st = Scene.load(path)
    st = dict(st or dict(width=256, height=256, iterations=16, depth=8, gamma=1.0,
                         ray_cast_epsilon=0.001, flip_y=0, seed=0))
    if spp is not None:
        st["iterations"] = spp
    if width is not None:
        st["width"] = width
    if height is not None:
        st["height"] = height
    arrays = Bvh(scene)
    dev = Device(device)
    ds = DeviceScene(dev, arrays)
    tgt = Target(ds, st["width"], st["height"], st["seed"])
    p = N.ConstantParams()
    p.camera = scene.camera()
    p.depth, p.width, p.height = st["depth"], st["width"], st["height"]
    p.flip_y = 1 if st["flip_y"] else 0
    p.inverted_gamma = 1.0 / st["gamma"]
    p.ray_cast_epsilon = st["ray_cast_epsilon"]
    p.textures_count = len(arrays.textures)
    p.current_iteration = 0.0
    tgt.render(p, st["iterations"])
    return tgt.framebuffer(), st["width"], st["height"], tgt.last_render_ms()


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("scene")
    ap.add_argument("out", help="output image, .pfm or .exr")
    ap.add_argument("--spp", type=int)
    ap.add_argument("--width", type=int)
    ap.add_argument("--height", type=int)
    ap.add_argument("--device", type=int, default=0)
    a = ap.parse_args(argv)
    rgba, w, h, ms = render(a.scene, a.spp, a.width, a.height, a.device)
    (write_exr if a.out.lower().endswith(".exr") else write_pfm)(a.out, rgba, w, h)
    print(f"{a.out}: {w}x{h}, {ms:.2f} ms", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
