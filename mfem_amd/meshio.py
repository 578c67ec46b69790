"""Mesh and matrix file formats used by the reference's scenes.

* Medit ``.mesh`` tetrahedral meshes (read via libigl's readMESH in the
  reference, src/simulation_state.cpp:62-68; 1-based indices).
* libigl ASCII ``.dmat`` matrices (``cols rows`` header, then column-major
  values; reference initial-state files scenes/data/*.dmat and
  src/simulation_state.cpp:263-308).
"""
from __future__ import annotations

import numpy as np


def read_mesh(path: str):
    """Read a Medit .mesh file. Returns V (nV x 3 float64), T (nT x 4 int32, 0-based)."""
    with open(path) as f:
        toks = []
        for line in f:
            line = line.split("#", 1)[0]
            toks.extend(line.split())
    i = 0
    V = T = None
    while i < len(toks):
        t = toks[i]
        if t == "Vertices":
            n = int(toks[i + 1])
            vals = np.array(toks[i + 2:i + 2 + 4 * n], dtype=np.float64).reshape(n, 4)
            V = vals[:, :3].copy()
            i += 2 + 4 * n
        elif t == "Tetrahedra":
            n = int(toks[i + 1])
            vals = np.array(toks[i + 2:i + 2 + 5 * n], dtype=np.int64).reshape(n, 5)
            T = (vals[:, :4] - 1).astype(np.int32)
            i += 2 + 5 * n
        elif t == "Triangles":
            n = int(toks[i + 1])
            i += 2 + 4 * n
        elif t == "Edges":
            n = int(toks[i + 1])
            i += 2 + 3 * n
        else:
            i += 1
    if V is None or T is None:
        raise ValueError(f"{path}: missing Vertices or Tetrahedra")
    return V, T


def read_obj(path: str):
    """Read an OBJ file as a triangle mesh, as the reference's 2D loader does
    with igl::read_triangle_mesh (src/simulation_state.cpp:46-60): "v" rows
    (z optional), "f" polygons fan-triangulated from their first corner,
    v/vt/vn and negative (relative) indices. Returns V (nV x 3), T (nT x 3)."""
    V, T = [], []
    with open(path) as f:
        for line in f:
            tok = line.split()
            if not tok:
                continue
            if tok[0] == "v":
                xyz = [float(t) for t in tok[1:4]]
                V.append(xyz + [0.0] * (3 - len(xyz)))
            elif tok[0] == "f":
                poly = []
                for t in tok[1:]:
                    k = int(t.split("/")[0])
                    if k == 0:
                        raise ValueError(f"{path}: bad face index")
                    poly.append(k - 1 if k > 0 else len(V) + k)
                for c in range(1, len(poly) - 1):
                    T.append([poly[0], poly[c], poly[c + 1]])
    V = np.array(V, dtype=np.float64).reshape(-1, 3)
    T = np.array(T, dtype=np.int32).reshape(-1, 3)
    if V.size == 0 or T.size == 0 or T.min() < 0 or T.max() >= V.shape[0]:
        raise ValueError(f"{path}: no vertices or faces, or an index out of range")
    return V, T


def remove_unreferenced(V: np.ndarray, T: np.ndarray):
    """Drop the vertices no element references, keeping the order of the
    others (igl::remove_unreferenced); returns (V, T renumbered)."""
    used = np.zeros(V.shape[0], dtype=bool)
    used[T.reshape(-1)] = True
    new_id = np.cumsum(used) - 1
    return V[used].copy(), new_id[T].astype(np.int32)


def write_obj(path: str, V: np.ndarray, T: np.ndarray) -> None:
    """Write vertices (nV x 2 or 3) and triangles as an OBJ file (1-based)."""
    with open(path, "w") as f:
        for v in np.asarray(V, dtype=np.float64):
            xyz = list(v) + [0.0] * (3 - len(v))
            f.write("v %.17g %.17g %.17g\n" % tuple(xyz[:3]))
        for t in np.asarray(T):
            f.write("f " + " ".join(str(int(i) + 1) for i in t) + "\n")


def write_mesh(path: str, V: np.ndarray, T: np.ndarray) -> None:
    """Write a Medit .mesh file (1-based indices, reference zero labels)."""
    with open(path, "w") as f:
        f.write("MeshVersionFormatted 1\nDimension 3\nVertices\n%d\n" % V.shape[0])
        for v in V:
            f.write("%.17g %.17g %.17g 0\n" % tuple(v))
        f.write("Triangles\n0\nTetrahedra\n%d\n" % T.shape[0])
        for t in T:
            f.write("%d %d %d %d 0\n" % tuple(int(i) + 1 for i in t))
        f.write("Edges\n0\nEnd\n")


def read_dmat(path: str) -> np.ndarray:
    """Read a libigl ASCII .dmat file into a (rows x cols) array."""
    with open(path) as f:
        head = f.readline().split()
        cols, rows = int(head[0]), int(head[1])
        data = np.array(f.read().split(), dtype=np.float64)
    if data.size != rows * cols:
        raise ValueError(f"{path}: expected {rows * cols} values, got {data.size}")
    return data.reshape(cols, rows).T.copy()


def write_dmat(path: str, A: np.ndarray) -> None:
    """Write a 2D array as a libigl ASCII .dmat file."""
    A = np.atleast_2d(np.asarray(A, dtype=np.float64))
    if A.ndim == 2 and A.shape[0] == 1 and A.shape[1] > 1:
        A = A.T
    rows, cols = A.shape
    with open(path, "w") as f:
        f.write(f"{cols} {rows}\n")
        for v in A.T.reshape(-1):
            f.write(f"{v:.17g}\n")
