"""glTF ingest (vlb_gltf_load) and the baked-light reader (vlb_gltf_read_bake), CPU.

Fixture: tests/golden/default_blender_cube.gltf, the reference's only asset
(src/vendor/default_blender_cube.gltf: a 2 m cube, 24 vertices, 36 uint16
indices, baseColorFactor 0.8).  Node transforms are checked against scipy's
quaternion rotation (an independent convention for glTF's [x, y, z, w]),
the scene bounds against the reference's rule (origin included, each
primitive's local corners through its own node matrix only).
"""
import base64
import json
import os
import struct

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import vlb

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CUBE = os.path.join(GOLDEN, "default_blender_cube.gltf")


def test_reference_cube_fixture():
    s, text = vlb.load_gltf(CUBE)
    assert s["idx"].shape == (12, 3) and s["xyz"].shape == (24, 3)
    assert np.array_equal(np.abs(s["xyz"]), np.ones((24, 3), np.float32))
    n = s["normals"]
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0) and np.allclose(np.abs(n).max(axis=1), 1.0)
    np.testing.assert_allclose(s["albedo"], 0.800000011920929, rtol=0, atol=0)
    assert np.array_equal(s["bounds_min"], [-1, -1, -1]) and np.array_equal(s["bounds_max"], [1, 1, 1])
    assert json.loads(text)["asset"]["generator"].startswith("Khronos glTF Blender")
    # closed: every edge is shared by exactly two triangles
    edges = {}
    for t in s["idx"]:
        for a, b in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0])):
            pa, pb = tuple(s["xyz"][a]), tuple(s["xyz"][b])
            key = (min(pa, pb), max(pa, pb))
            edges[key] = edges.get(key, 0) + 1
    assert set(edges.values()) == {2}


def _tri_gltf(tmp_path, nodes, scene_roots, index_type=5123, glb=False, color=None):
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    nrm = np.array([[0, 0, 2.0]] * 3, np.float32)        # unnormalised on purpose
    ind = np.array([0, 1, 2], {5121: np.uint8, 5123: np.uint16, 5125: np.uint32}[index_type])
    blob = pos.tobytes() + nrm.tobytes() + ind.tobytes()
    blob += b"\0" * (-len(blob) % 4)
    doc = {
        "asset": {"version": "2.0"},
        "scenes": [{"nodes": scene_roots}],
        "nodes": nodes,
        "meshes": [{"primitives": [{"attributes": {"POSITION": 0, "NORMAL": 1}, "indices": 2,
                                     "material": 0}]}],
        "materials": [{"pbrMetallicRoughness": {"baseColorFactor": color or [0.25, 0.5, 0.75, 1.0]}}],
        "accessors": [
            {"bufferView": 0, "componentType": 5126, "count": 3, "type": "VEC3"},
            {"bufferView": 1, "componentType": 5126, "count": 3, "type": "VEC3"},
            {"bufferView": 2, "componentType": index_type, "count": 3, "type": "SCALAR"},
        ],
        "bufferViews": [
            {"buffer": 0, "byteOffset": 0, "byteLength": 36},
            {"buffer": 0, "byteOffset": 36, "byteLength": 36},
            {"buffer": 0, "byteOffset": 72, "byteLength": ind.nbytes},
        ],
        "buffers": [{"byteLength": len(blob)}],
    }
    if glb:
        js = json.dumps(doc).encode()
        js += b" " * (-len(js) % 4)
        body = struct.pack("<II", len(js), 0x4E4F534A) + js + struct.pack("<II", len(blob), 0x004E4942) + blob
        data = struct.pack("<III", 0x46546C67, 2, 12 + len(body)) + body
        path = tmp_path / "tri.glb"
        path.write_bytes(data)
    else:
        doc["buffers"][0]["uri"] = "data:application/octet-stream;base64," + base64.b64encode(blob).decode()
        path = tmp_path / "tri.gltf"
        path.write_text(json.dumps(doc))
    return str(path), pos, nrm


def _trs(node):
    M = np.eye(4)
    if "matrix" in node:
        M = np.array(node["matrix"], np.float64).reshape(4, 4).T
    T, R, S = np.eye(4), np.eye(4), np.eye(4)
    if "translation" in node:
        T[:3, 3] = node["translation"]
    if "rotation" in node:
        R[:3, :3] = Rotation.from_quat(node["rotation"]).as_matrix()   # scipy: (x, y, z, w)
    if "scale" in node:
        S[:3, :3] = np.diag(node["scale"])
    return T @ R @ S @ M


@pytest.mark.parametrize("index_type", [5121, 5123, 5125])
@pytest.mark.parametrize("glb", [False, True])
def test_node_transforms_normals_and_bounds(tmp_path, index_type, glb):
    q = Rotation.from_euler("xyz", [30, -50, 80], degrees=True).as_quat()
    child = {"mesh": 0, "translation": [0.5, -2.0, 3.0], "rotation": q.tolist(), "scale": [2.0, 1.0, 0.5]}
    parent = {"children": [1], "translation": [10.0, 0.0, 0.0],
              "matrix": [1, 0, 0, 0, 0, 2, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1]}
    path, pos, nrm = _tri_gltf(tmp_path, [parent, child], [0], index_type, glb)
    s, _ = vlb.load_gltf(path)
    W = _trs(parent) @ _trs(child)
    ref = (W[:3, :3] @ pos.T.astype(np.float64)).T + W[:3, 3]
    np.testing.assert_allclose(s["xyz"], ref, rtol=1e-6, atol=1e-6)
    nref = (np.linalg.inv(W[:3, :3]).T @ (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).T).T
    np.testing.assert_allclose(s["normals"], nref, rtol=1e-6, atol=1e-6)
    assert s["idx"].tolist() == [[0, 1, 2]]
    np.testing.assert_allclose(s["albedo"], [[0.25, 0.5, 0.75]])
    # bounds: origin included; local min/max corners through the child's own matrix
    L = _trs(child)
    c0 = L[:3, :3] @ pos.min(axis=0) + L[:3, 3]
    c1 = L[:3, :3] @ pos.max(axis=0) + L[:3, 3]
    np.testing.assert_allclose(s["bounds_min"], np.minimum(0.0, c0), rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(s["bounds_max"], np.maximum(0.0, c1), rtol=1e-6, atol=1e-6)


def test_black_or_missing_base_colour_is_white(tmp_path):
    path, _, _ = _tri_gltf(tmp_path, [{"mesh": 0}], [0], color=[0, 0, 0, 0])
    s, _ = vlb.load_gltf(path)
    np.testing.assert_array_equal(s["albedo"], [[1.0, 1.0, 1.0]])


def test_load_errors(tmp_path):
    with pytest.raises(vlb.VlbError) as e:
        vlb.load_gltf(str(tmp_path / "missing.gltf"))
    assert e.value.status == vlb.ERR_NOT_FOUND
    bad = tmp_path / "scene.obj"
    bad.write_text("v 0 0 0")
    with pytest.raises(vlb.VlbError) as e:
        vlb.load_gltf(str(bad))
    assert e.value.status == vlb.ERR_UNSUPPORTED
    broken = tmp_path / "broken.gltf"
    broken.write_text("{ not json")
    with pytest.raises(vlb.VlbError):
        vlb.load_gltf(str(broken))


def test_baked_light_round_trip():
    text = open(CUBE).read()
    rng = np.random.default_rng(4)
    coeffs = rng.normal(size=(343, 16, 3)).astype(np.float32)
    step = np.array([0.3333333, 0.25, 0.5], np.float32)
    baked = vlb.gltf_bake_json(text, coeffs, step)
    got, gstep, k = vlb.read_baked_gltf(baked)
    assert k == 16
    assert np.array_equal(got.view(np.uint32), coeffs.ravel().view(np.uint32))
    assert np.array_equal(gstep, step)
    with pytest.raises(vlb.VlbError) as e:
        vlb.read_baked_gltf(text)           # no "light" in the source asset
    assert e.value.status == vlb.ERR_NOT_FOUND
