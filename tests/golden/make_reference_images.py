"""Writes tests/golden/cornell_64.orns and cornell_64_4096spp.pfm: a scene
file and its 4,096-spp reference image, rendered by the oracle (the C
restatement of the WGSL megakernel) on the CPU.

The image is the oracle's post-processed framebuffer after one render() of
4,096 iterations from the seeded state (RNG = pixel index, accumulation 0,
current_iteration 0), stored as PFM (RGB, row 0 = bottom). The GPU test
renders the same file on the device and must reproduce these bits; the CPU
test re-renders it with the oracle. Run from the repository root:
    python tests/golden/make_reference_images.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle_lib as O  # noqa: E402
import zig_ornament_amd as zo  # noqa: E402
from zig_ornament_amd import workloads as W  # noqa: E402

SCENE = os.path.join(HERE, "cornell_64.orns")
IMAGE = os.path.join(HERE, "cornell_64_4096spp.pfm")


def oracle_image(path, nthreads=8):
    """Oracle render of a scene file at its stored settings."""
    spec, st = W.load_spec(path)
    arrays = O.build_bvh(spec)
    n = spec.width * spec.height
    p = O.constant_params(spec)
    acc, rng = np.zeros((n, 4), np.float32), O.seed_rng(n, spec.seed)
    fb, _ = O.OracleScene(arrays).render(p, spec.spp, acc, rng, nthreads=nthreads)
    return fb, spec


def main():
    O.build()
    spec = W.cornell(width=64, height=64, spp=4096)
    W.save_spec(spec, SCENE)
    fb, spec = oracle_image(SCENE)
    zo.write_pfm(IMAGE, fb, spec.width, spec.height)
    print(SCENE, os.path.getsize(SCENE), IMAGE, os.path.getsize(IMAGE))


if __name__ == "__main__":
    main()
