"""Writes the mesh-ingestion fixtures.

Fixtures (tests/golden/meshes/): small glTF 2.0 files built here with known
content (node TRS and matrix transforms, a child node, interleaved vertex
data, uint8 / uint16 / uint32 indices, a primitive without indices, a LINES
primitive that must be skipped, a Draco primitive that must be skipped), and a
reduced synthetic vehicle laid with flow along y and up along x, as GLB and as
binary STL.  The tests check them against closed forms, and against the
reference's own parse.js / prepare.js through make_prepare_golden.py.

usage: python tests/golden/make_mesh_golden.py
"""
from __future__ import annotations

import base64
import json
import struct
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
OUT = HERE / "meshes"
sys.path.insert(0, str(HERE.parent.parent))


def _pad4(b: bytes, fill: bytes = b"\0") -> bytes:
    return b + fill * ((4 - len(b) % 4) % 4)


def glb(doc: dict, binary: bytes) -> bytes:
    j = _pad4(json.dumps(doc, separators=(",", ":")).encode(), b" ")
    b = _pad4(binary)
    total = 12 + 8 + len(j) + 8 + len(b)
    return (struct.pack("<III", 0x46546C67, 2, total) + struct.pack("<II", len(j), 0x4E4F534A) + j
            + struct.pack("<II", len(b), 0x004E4942) + b)


def nodes_case() -> tuple[dict, bytes]:
    """Two tetrahedra and a triangle under a node hierarchy."""
    tet = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32)
    ind16 = np.array([0, 2, 1, 0, 1, 3, 0, 3, 2, 1, 2, 3], np.uint16)
    # interleaved position + normal (stride 24) for the second tetrahedron
    inter = np.zeros((4, 6), np.float32)
    inter[:, :3] = tet * 2.0 + np.array([3, 0, 0], np.float32)
    inter[:, 3:] = 0.5
    ind8 = np.array([0, 1, 2, 0, 3, 1, 0, 2, 3, 1, 3, 2], np.uint8)
    tri = np.array([[0, 0, 0], [0.5, 0.25, 0], [0, 0.75, 0.125]], np.float32)   # no indices
    lines = np.array([[9, 9, 9], [8, 8, 8]], np.float32)
    ind32 = np.array([0, 1, 2, 1, 2, 3], np.uint32)
    parts = [tet.tobytes(), ind16.tobytes(), inter.tobytes(), ind8.tobytes(), tri.tobytes(),
             lines.tobytes(), (tet * 4).tobytes(), ind32.tobytes()]
    views, off, binary = [], 0, b""
    for k, p in enumerate(parts):
        p = _pad4(p)
        v = {"buffer": 0, "byteOffset": off, "byteLength": len(p)}
        if k == 2:
            v["byteStride"] = 24
        views.append(v)
        binary += p
        off += len(p)
    acc = [
        {"bufferView": 0, "componentType": 5126, "count": 4, "type": "VEC3"},
        {"bufferView": 1, "componentType": 5123, "count": 12, "type": "SCALAR"},
        {"bufferView": 2, "componentType": 5126, "count": 4, "type": "VEC3"},
        {"bufferView": 3, "componentType": 5121, "count": 12, "type": "SCALAR"},
        {"bufferView": 4, "componentType": 5126, "count": 3, "type": "VEC3"},
        {"bufferView": 5, "componentType": 5126, "count": 2, "type": "VEC3"},
        {"bufferView": 6, "componentType": 5126, "count": 4, "type": "VEC3"},
        {"bufferView": 7, "componentType": 5125, "count": 6, "type": "SCALAR"},
    ]
    meshes = [
        {"primitives": [{"attributes": {"POSITION": 0}, "indices": 1},
                        {"attributes": {"POSITION": 5}, "mode": 1}]},
        {"primitives": [{"attributes": {"POSITION": 2, "NORMAL": 2}, "indices": 3},
                        {"attributes": {"POSITION": 4}}]},
        {"primitives": [{"attributes": {"POSITION": 6}, "indices": 7,
                         "extensions": {"KHR_draco_mesh_compression": {}}},
                        {"attributes": {"POSITION": 6}, "indices": 7, "mode": 4}]},
    ]
    nodes = [
        {"mesh": 0, "translation": [10, 20, 30], "rotation": [0, 0, 0.38268343, 0.92387953],
         "scale": [2, 2, 2], "children": [1]},
        {"mesh": 1, "matrix": [0, 1, 0, 0, -1, 0, 0, 0, 0, 0, 1, 0, 5, 6, 7, 1]},
        {"mesh": 2, "rotation": [0.1, 0.2, 0.3, 0.9273618495495703], "translation": [-1, 2, -3]},
    ]
    doc = {"asset": {"version": "2.0", "generator": "fixture"}, "scene": 0,
           "scenes": [{"name": "nodes", "nodes": [0, 2]}], "nodes": nodes, "meshes": meshes,
           "accessors": acc, "bufferViews": views, "buffers": [{"byteLength": len(binary)}]}
    return doc, binary


def vehicle_small(seed=3):
    from cfd__amd import meshes
    pos, idx, _ = meshes.vehicle(seed=seed, nu=60, nv=56, wheel_segments=200)
    p = pos.reshape(-1, 3).astype(np.float64)
    # file frame: flow along y (reversed: nose towards +y), up along x, in mm
    q = np.stack([p[:, 2], -p[:, 0], p[:, 1]], axis=1) * 45.0
    return q.astype(np.float32), idx.astype(np.uint32)


def stl_binary(pos: np.ndarray, idx: np.ndarray, header=b"fixture") -> bytes:
    t = pos.reshape(-1, 3)[idx.reshape(-1, 3)]                   # (n, 3, 3)
    rec = np.zeros((t.shape[0], 50), np.uint8)
    rec[:, 12:48] = np.ascontiguousarray(t.astype("<f4").reshape(t.shape[0], 9)).view(np.uint8)
    return header.ljust(80, b" ") + struct.pack("<I", t.shape[0]) + rec.tobytes()


def main():
    OUT.mkdir(exist_ok=True)
    doc, binary = nodes_case()
    (OUT / "nodes.glb").write_bytes(glb(doc, binary))
    d2 = json.loads(json.dumps(doc))
    d2["buffers"] = [{"byteLength": len(binary),
                      "uri": "data:application/octet-stream;base64," +
                             base64.b64encode(binary).decode()}]
    (OUT / "nodes.gltf").write_text(json.dumps(d2))
    vp, vi = vehicle_small()
    bv = vp.tobytes()
    bi = vi.tobytes()
    vdoc = {"asset": {"version": "2.0"}, "nodes": [{"mesh": 0}],
            "meshes": [{"primitives": [{"attributes": {"POSITION": 0}, "indices": 1}]}],
            "accessors": [{"bufferView": 0, "componentType": 5126, "count": len(vp),
                           "type": "VEC3"},
                          {"bufferView": 1, "componentType": 5125, "count": len(vi),
                           "type": "SCALAR"}],
            "bufferViews": [{"buffer": 0, "byteOffset": 0, "byteLength": len(bv)},
                            {"buffer": 0, "byteOffset": len(bv), "byteLength": len(bi)}],
            "buffers": [{"byteLength": len(bv) + len(bi)}]}
    (OUT / "vehicle.glb").write_bytes(glb(vdoc, bv + bi))
    (OUT / "vehicle.stl").write_bytes(stl_binary(vp, vi))
    print(sorted(p.name for p in OUT.iterdir()))


if __name__ == "__main__":
    main()
