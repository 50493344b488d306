"""Host-side logic and the C ABI on the CPU (no GPU calls): the library loads
and exports every symbol of include/vlb_bake.h, contexts fail loudly without a
device, and the host helpers match the reference's algorithms."""
import base64
import ctypes
import json
import os
import re
import subprocess

import numpy as np
import pytest

import orc
import vlb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_symbols():
    text = open(os.path.join(ROOT, "include", "vlb_bake.h")).read()
    decl = r"(?:vlb_status|int32_t|void|const char\s*\*)\s+(vlb_[a-z0-9_]+)\s*\("
    return sorted(set(re.findall(decl, text)))


def test_library_exports_every_declared_symbol():
    syms = _header_symbols()
    assert len(syms) >= 40
    missing = [s for s in syms if not hasattr(vlb.lib, s)]
    assert not missing, missing
    assert vlb.lib.vlb_abi_version() == 1


def test_context_fails_loudly_without_device():
    import glob
    if glob.glob("/dev/kfd"):
        pytest.skip("a GPU is visible")
    with pytest.raises(vlb.VlbError) as e:
        vlb.Context(0)
    assert e.value.status == vlb.ERR_NO_DEVICE


def test_failure_path_entry_points_reject_bad_handles():
    """Host-side argument checks of the failure-handling entry points: no
    device call is made for a null handle or a bad site / deadline."""
    L = vlb.lib
    assert L.vlb_debug_inject(None, vlb.INJECT_PLAN_EXECUTE, 1) == vlb.ERR_INVALID
    assert L.vlb_cancel(None, 1) == vlb.ERR_INVALID
    assert L.vlb_server_hold(None, 1) == vlb.ERR_INVALID
    assert L.vlb_wait_for(None, 1, 10.0, None, 0, None) == vlb.ERR_INVALID
    assert L.vlb_batch_set_timing(None, 1) == vlb.ERR_INVALID
    assert L.vlb_scene_release(None, 1) == vlb.ERR_INVALID
    assert vlb.STATUS_NAMES[vlb.ERR_TIMEOUT] == "TIMEOUT"
    assert vlb.STATUS_NAMES[vlb.ERR_CANCELLED] == "CANCELLED"
    assert vlb.STATUS_NAMES[vlb.ERR_INJECTED] == "INJECTED"


def test_probe_grid_reference_bfs_order():
    # light_baker.cpp:80-101 with probesCount3D = 7: first 7 probes walk x,
    # the next 42 walk y = 1..6 for each x, the rest walk z = 1..6 for each
    # of the first 49; positions are repeated float += step.
    b0 = np.array([-1.5, 0.25, -3.0], np.float32)
    b1 = np.array([4.0, 7.1, 2.2], np.float32)
    pos, step = vlb.probe_grid(b0, b1, [7, 7, 7], vlb.ORDER_REFERENCE_BFS)
    step_ref = (b1 - b0) / (np.float32(7) - np.float32(1))
    assert np.array_equal(step, step_ref)
    expect = [b0.copy()]
    for dim in range(3):
        for q in list(expect):
            p = q.copy()
            for _ in range(6):
                p[dim] = np.float32(p[dim] + step[dim])
                expect.append(p.copy())
    expect = np.array(expect, np.float32)
    assert pos.shape == (343, 3) and np.array_equal(pos, expect)
    # explicit index structure
    ijk = np.rint((pos - b0) / step).astype(int)
    assert [tuple(v) for v in ijk[:7]] == [(i, 0, 0) for i in range(7)]
    assert [tuple(v) for v in ijk[7:13]] == [(0, j, 0) for j in range(1, 7)]
    assert tuple(ijk[49]) == (0, 0, 1) and tuple(ijk[342]) == (6, 6, 6)
    o, _ = orc.probe_grid(b0, b1, [7, 7, 7], vlb.ORDER_REFERENCE_BFS)
    assert np.array_equal(o, pos)


def test_probe_grid_linear_order():
    pos, step = vlb.probe_grid([0, 0, 0], [3, 2, 1], [4, 3, 2], vlb.ORDER_LINEAR)
    assert pos.shape == (24, 3)
    np.testing.assert_array_equal(pos[1], [1, 0, 0])
    np.testing.assert_array_equal(pos[4], [0, 1, 0])
    np.testing.assert_array_equal(pos[12], [0, 0, 1])
    o, _ = orc.probe_grid([0, 0, 0], [3, 2, 1], [4, 3, 2], vlb.ORDER_LINEAR)
    assert np.array_equal(o, pos)


@pytest.mark.parametrize("data", [b"", b"f", b"fo", b"foo", b"foob", b"fooba", b"foobar"])
def test_base64_rfc4648_vectors(data):
    assert vlb.base64_encode(data) == base64.b64encode(data).decode()
    assert orc.base64(data) == base64.b64encode(data).decode()


def test_base64_random_bytes():
    rng = np.random.default_rng(0)
    for n in (1, 2, 3, 191, 192, 193, 65856):
        b = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        assert vlb.base64_encode(b) == base64.b64encode(b).decode()


def test_gltf_bake_json_structure():
    src = json.dumps({"asset": {"version": "2.0", "generator": "t"}, "scene": 0,
                      "buffers": [{"byteLength": 4, "uri": "data:application/octet-stream;base64,AAAAAA=="}],
                      "bufferViews": [{"buffer": 0, "byteLength": 4}],
                      "nodes": [{"name": "Cube", "rotation": [-0.7071067690849304, 0, 0, 0.7071067690849304]}]})
    rng = np.random.default_rng(1)
    coeffs = rng.normal(size=(343, 16, 3)).astype(np.float32)
    step = np.array([1.0, 0.5, 1.6666666269302368], np.float32)
    out = vlb.gltf_bake_json(src, coeffs, step)
    j = json.loads(out)
    assert len(j["buffers"]) == 2 and len(j["bufferViews"]) == 2
    assert j["light"] == {"bufferView": 1, "gridStep": {"x": 1.0, "y": 0.5, "z": float(step[2])}, "lmax": 16}
    bv = j["bufferViews"][1]
    assert bv == {"buffer": 1, "byteLength": 343 * 16 * 12, "byteOffset": 0}
    uri = j["buffers"][1]["uri"]
    assert uri.startswith("data:application/octet-stream;base64,")
    raw = base64.b64decode(uri.split(",", 1)[1])
    assert np.array_equal(np.frombuffer(raw, np.float32).reshape(343, 16, 3), coeffs)
    assert j["nodes"] == json.loads(src)["nodes"]
    # nlohmann dump(4) layout: sorted keys, 4-space indent, ": " separator
    assert out == json.dumps(j, indent=4, sort_keys=True) + "\n"


def test_gltf_bake_json_rejects_bad_input():
    with pytest.raises(vlb.VlbError):
        vlb.gltf_bake_json("{not json", np.zeros((1, 9, 3), np.float32), [1, 1, 1])
    with pytest.raises(vlb.VlbError):
        vlb.gltf_bake_json("[1, 2]", np.zeros((1, 9, 3), np.float32), [1, 1, 1])


def test_cxx_facade_reports_missing_device():
    import glob
    if glob.glob("/dev/kfd"):
        pytest.skip("a GPU is visible")
    exe = os.path.join(ROOT, "vulkan-light-bakery_amd", "baker")
    if not os.path.exists(exe):
        pytest.skip("baker CLI not built")
    r = subprocess.run([exe, "room:2"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "status 11" in r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "Select scene" in r.stderr


def test_bench_refuses_more_than_one_gpu():
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    for argv, env in ((["--gpus", "2"], {}), ([], {"WORLD_SIZE": "4"})):
        r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), *argv], cwd=root,
                           capture_output=True, text=True, timeout=120, env={**os.environ, **env})
        assert r.returncode == 2
        assert "one GPU" in r.stderr and r.stdout == ""
