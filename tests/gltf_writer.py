"""A small glTF 2.0 writer for the loader tests: builds the JSON and one
binary buffer from numpy arrays and writes it as a .gltf with a base64 data
URI, as a .gltf beside a .bin file, or as a binary .glb."""
from __future__ import annotations

import base64
import json
import struct
import urllib.parse
from pathlib import Path

import numpy as np

FLOAT, U8, U16, U32 = 5126, 5121, 5123, 5125
_COMP = {np.dtype(np.float32): FLOAT, np.dtype(np.uint8): U8, np.dtype(np.uint16): U16,
         np.dtype(np.uint32): U32}


class GltfBuilder:
    def __init__(self):
        self.bin = bytearray()
        self.doc = {"asset": {"version": "2.0"}, "buffers": [], "bufferViews": [], "accessors": [],
                    "meshes": [], "nodes": [], "scenes": []}

    def _view(self, data: bytes, stride: int | None = None) -> int:
        while len(self.bin) % 4:
            self.bin.append(0)
        v = {"buffer": 0, "byteOffset": len(self.bin), "byteLength": len(data)}
        if stride:
            v["byteStride"] = stride
        self.bin += data
        self.doc["bufferViews"].append(v)
        return len(self.doc["bufferViews"]) - 1

    def _accessor(self, view: int, dtype, count: int, type_: str, offset: int = 0,
                  minmax: np.ndarray | None = None) -> int:
        a = {"bufferView": view, "componentType": _COMP[np.dtype(dtype)], "count": int(count),
             "type": type_}
        if offset:
            a["byteOffset"] = offset
        if minmax is not None and len(minmax):
            a["min"] = [float(x) for x in minmax.min(axis=0)]
            a["max"] = [float(x) for x in minmax.max(axis=0)]
        self.doc["accessors"].append(a)
        return len(self.doc["accessors"]) - 1

    def vec3(self, arr, with_minmax: bool = True) -> int:
        arr = np.ascontiguousarray(arr, np.float32).reshape(-1, 3)
        return self._accessor(self._view(arr.tobytes()), np.float32, len(arr), "VEC3",
                              minmax=arr if with_minmax else None)

    def interleaved(self, pos, nrm) -> tuple[int, int]:
        """Positions and normals in one view with byteStride 24."""
        pos = np.asarray(pos, np.float32).reshape(-1, 3)
        nrm = np.asarray(nrm, np.float32).reshape(-1, 3)
        view = self._view(np.concatenate([pos, nrm], axis=1).tobytes(), stride=24)
        return (self._accessor(view, np.float32, len(pos), "VEC3", 0, pos),
                self._accessor(view, np.float32, len(pos), "VEC3", 12))

    def indices(self, idx, dtype) -> int:
        idx = np.ascontiguousarray(idx, dtype).reshape(-1)
        if len(idx) == 0:
            self.doc["accessors"].append({"componentType": _COMP[np.dtype(dtype)], "count": 0,
                                          "type": "SCALAR"})
            return len(self.doc["accessors"]) - 1
        return self._accessor(self._view(idx.tobytes()), dtype, len(idx), "SCALAR")

    def mesh(self, primitives: list[dict]) -> int:
        self.doc["meshes"].append({"primitives": primitives})
        return len(self.doc["meshes"]) - 1

    def node(self, **kw) -> int:
        self.doc["nodes"].append(kw)
        return len(self.doc["nodes"]) - 1

    def write(self, path: Path, mode: str = "data") -> Path:
        """mode: "data" (base64 URI), "file" (uri of a .bin beside it, with a
        space in its name) or "glb"."""
        path = Path(path)
        doc = json.loads(json.dumps(self.doc))
        data = bytes(self.bin)
        if mode == "data":
            doc["buffers"] = [{"byteLength": len(data),
                               "uri": "data:application/octet-stream;base64," +
                                      base64.b64encode(data).decode()}]
            path.write_text(json.dumps(doc))
        elif mode == "file":
            name = path.stem + " buffer.bin"
            (path.parent / name).write_bytes(data)
            doc["buffers"] = [{"byteLength": len(data), "uri": urllib.parse.quote(name)}]
            path.write_text(json.dumps(doc))
        elif mode == "glb":
            doc["buffers"] = [{"byteLength": len(data)}]
            js = json.dumps(doc).encode()
            js += b" " * (-len(js) % 4)
            bn = data + b"\0" * (-len(data) % 4)
            total = 12 + 8 + len(js) + 8 + len(bn)
            out = struct.pack("<III", 0x46546C67, 2, total)
            out += struct.pack("<II", len(js), 0x4E4F534A) + js
            out += struct.pack("<II", len(bn), 0x004E4942) + bn
            path.write_bytes(out)
        else:
            raise ValueError(mode)
        return path


def export_host_scene(hs, path: Path, mode: str = "data") -> Path:
    """A HostScene as glTF: one primitive per geometry with the vertices its
    triangles use, one mesh, one node without a transform, and one glTF
    material per scene material (each geometry must use a single one)."""
    b = GltfBuilder()
    prims = []
    gf = hs.geom_first
    for g in range(len(gf) - 1):
        tri = hs.indices[gf[g]:gf[g + 1]]
        used, inv = np.unique(tri.reshape(-1), return_inverse=True)
        mats = np.unique(hs.tri_material[gf[g]:gf[g + 1]])
        assert len(mats) == 1, "geometry with several materials"
        prims.append({"attributes": {"POSITION": b.vec3(hs.vertices[used, :3]),
                                     "NORMAL": b.vec3(hs.vertices[used, 3:], False)},
                      "indices": b.indices(inv.astype(np.uint32), np.uint32),
                      "material": int(mats[0])})
    b.doc["materials"] = [{} for _ in range(len(hs.albedo))]
    m = b.mesh(prims)
    b.doc["scenes"] = [{"nodes": [b.node(mesh=m)]}]
    b.doc["scene"] = 0
    return b.write(path, mode)
