"""GPU: the committed scene file renders on the device to the committed
4,096-spp reference image, bit for bit (the image was made by the oracle
on the CPU from the same file)."""
import os

import numpy as np
import pytest

import zig_ornament_amd as zo
from zig_ornament_amd import render_file

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_scene_file_renders_the_reference_image(device, tmp_path):
    rgba, w, h, ms = render_file.render(os.path.join(GOLDEN, "cornell_64.orns"))
    print(f"[files] cornell 64x64 4096 spp: {ms:.2f} ms", flush=True)
    ref, rw, rh = zo.read_pfm(os.path.join(GOLDEN, "cornell_64_4096spp.pfm"))
    assert (w, h) == (rw, rh)
    assert np.array_equal(rgba[:, :3], ref[:, :3])
    out = str(tmp_path / "again.pfm")
    assert render_file.main([os.path.join(GOLDEN, "cornell_64.orns"), out]) == 0
    again, _, _ = zo.read_pfm(out)
    assert np.array_equal(again, ref)
