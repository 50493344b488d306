"""End-to-end drop-in check on the reference's own asset: the baker CLI (the
reference main's shape: LightBaker(name), bake(), serialize()) bakes
default_blender_cube.gltf with the reference defaults (7x7x7 breadth-first
grid over the scene bounds, env_map.rchit shading with sRGB + RGBA8, 16
literal-constant coefficients) at a reduced equirect grid; the written
baked glTF is read back and compared with the FP64 oracle: <= 1e-5 relative.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import orc
import vlb
from conftest import rel_err

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CUBE = os.path.join(ROOT, "tests", "golden", "default_blender_cube.gltf")


def test_cli_bakes_the_reference_cube_asset(tmp_path):
    shutil.copy(CUBE, tmp_path / "default_blender_cube.gltf")
    exe = os.path.join(ROOT, "vulkan-light-bakery_amd", "baker")
    r = subprocess.run([exe, "default_blender_cube.gltf", "--equirect", "314", "100"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    coeffs, step, k = vlb.read_baked_gltf((tmp_path / "baked_default_blender_cube.gltf").read_text())
    assert k == 16 and coeffs.size == 343 * 16 * 3
    coeffs = coeffs.reshape(343, 16, 3)
    s, _ = vlb.load_gltf(CUBE)
    probes, gstep = vlb.probe_grid(s["bounds_min"], s["bounds_max"], [7, 7, 7], vlb.ORDER_REFERENCE_BFS)
    assert np.array_equal(step, gstep)
    p = vlb.Params(probes=probes, dir_mode=vlb.DIRS_EQUIRECT, equirect_w=314, equirect_h=100, sh_order=3,
                   sh_constants=vlb.SH_LITERAL, shade_mode=vlb.SHADE_REFERENCE,
                   ref_light=(1.0, 10.0, 1.0), ref_bias=0.005, ref_flags=vlb.REF_SRGB | vlb.REF_UNORM8)
    ref = orc.Scene(s["xyz"], s["idx"], normals=s["normals"], albedo=s["albedo"]).bake(p, mode64=True)["coeffs"]
    err = rel_err(coeffs, ref)
    assert err.max() <= 1e-5, err.max()
