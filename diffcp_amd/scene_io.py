"""Triangle-mesh PLY reading and writing (the formats Open3D reads and writes
for the reference: ``o3d.io.read_triangle_mesh`` in initialization.py:105 and
``o3d.io.write_triangle_mesh`` in simulation_without_grad.py:520).

Supported: ``ascii`` and ``binary_little_endian`` 1.0, a ``vertex`` element
whose first three properties are x, y, z (float or double; further vertex
properties are skipped) and a ``face`` element with one ``list`` property of
vertex indices (triangles only).
"""
from __future__ import annotations

import os
from typing import List, Tuple

import numpy as np

_TYPES = {"char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1", "short": "i2", "int16": "i2",
          "ushort": "u2", "uint16": "u2", "int": "i4", "int32": "i4", "uint": "u4", "uint32": "u4",
          "float": "f4", "float32": "f4", "double": "f8", "float64": "f8"}


def _header(f) -> Tuple[str, List[Tuple[str, int, List[List[str]]]]]:
    if f.readline().strip() != b"ply":
        raise ValueError("not a PLY file")
    fmt, elements = None, []
    while True:
        line = f.readline()
        if not line:
            raise ValueError("PLY header without end_header")
        tok = line.decode("ascii", "replace").split()
        if not tok or tok[0] in ("comment", "obj_info"):
            continue
        if tok[0] == "format":
            fmt = tok[1]
        elif tok[0] == "element":
            elements.append((tok[1], int(tok[2]), []))
        elif tok[0] == "property":
            elements[-1][2].append(tok[1:])
        elif tok[0] == "end_header":
            break
    if fmt not in ("ascii", "binary_little_endian"):
        raise ValueError(f"unsupported PLY format {fmt}")
    return fmt, elements


def read_ply(path: str) -> Tuple[np.ndarray, np.ndarray]:
    """(vertices [N, 3] float64, triangles [M, 3] int32) of a triangle-mesh PLY."""
    with open(path, "rb") as f:
        fmt, elements = _header(f)
        verts, tris = None, None
        if fmt == "ascii":
            rest = f.read().decode("ascii").split()
            pos = 0
            for name, n, props in elements:
                if name == "vertex":
                    k = len(props)
                    a = np.array(rest[pos:pos + n * k], dtype=np.float64).reshape(n, k)
                    verts = a[:, :3]
                    pos += n * k
                elif name == "face":
                    out = np.empty((n, 3), np.int64)
                    for i in range(n):
                        c = int(rest[pos])
                        if c != 3:
                            raise ValueError("only triangle faces are supported")
                        out[i] = [int(v) for v in rest[pos + 1:pos + 4]]
                        pos += 1 + c
                    tris = out
                else:
                    pos += n * len(props)
        else:
            for name, n, props in elements:
                if name == "vertex":
                    dt = np.dtype([(f"p{i}", "<" + _TYPES[p[0]]) for i, p in enumerate(props)])
                    a = np.frombuffer(f.read(dt.itemsize * n), dtype=dt, count=n)
                    verts = np.stack([a["p0"], a["p1"], a["p2"]], axis=1).astype(np.float64)
                elif name == "face":
                    if len(props) != 1 or props[0][0] != "list":
                        raise ValueError("face element must hold one list property")
                    ct, it = "<" + _TYPES[props[0][1]], "<" + _TYPES[props[0][2]]
                    # every face a triangle: fixed-size records
                    dt = np.dtype([("c", ct), ("v", it, (3,))])
                    a = np.frombuffer(f.read(dt.itemsize * n), dtype=dt, count=n)
                    if np.any(a["c"] != 3):
                        raise ValueError("only triangle faces are supported")
                    tris = a["v"].astype(np.int64)
                else:
                    raise ValueError(f"unsupported binary PLY element {name}")
    if verts is None or tris is None:
        raise ValueError("PLY without vertex or face element")
    return np.ascontiguousarray(verts), np.ascontiguousarray(tris, dtype=np.int32)


def write_ply(path: str, vertices: np.ndarray, triangles: np.ndarray) -> None:
    """Binary little-endian triangle-mesh PLY (double vertices, uchar/uint faces)."""
    v = np.ascontiguousarray(np.asarray(vertices, np.float64).reshape(-1, 3))
    t = np.asarray(triangles).reshape(-1, 3)
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    head = ("ply\nformat binary_little_endian 1.0\n"
            f"element vertex {len(v)}\nproperty double x\nproperty double y\nproperty double z\n"
            f"element face {len(t)}\nproperty list uchar uint vertex_indices\nend_header\n")
    faces = np.empty(len(t), dtype=np.dtype([("c", "u1"), ("v", "<u4", (3,))]))
    faces["c"] = 3
    faces["v"] = t
    with open(path, "wb") as f:
        f.write(head.encode("ascii"))
        f.write(v.astype("<f8").tobytes())
        f.write(faces.tobytes())
