"""CPU-side checks of the C ABI library and the host-side scene tooling (no GPU)."""
import os
import re

import numpy as np
import pytest

from mfem_amd import meshio, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "mfem_hip.h")).read()
    return sorted(set(re.findall(r"\b(mfem_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from mfem_amd import native
    lib = native.lib()
    names = declared_symbols()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(native.EXPORTED)
    assert lib.mfem_version() == 1


def test_library_is_gfx950_code_object():
    """The shared object carries a gfx950 offload bundle (built for MI355X)."""
    from mfem_amd import native
    blob = open(native.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_splitmix_stream_is_reproducible():
    a = scenes.splitmix64_uniform(123, 1, 1000)
    b = scenes.splitmix64_uniform(123, 1, 1000)
    c = scenes.splitmix64_uniform(124, 1, 1000)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert 0.0 <= a.min() and a.max() < 1.0 and abs(a.mean() - 0.5) < 0.05


def test_config_sizes_match_the_kuhn_counts():
    """Vertex / tet / free-DOF counts of the benchmark configurations."""
    for (nx, ny, nz), nT, nV, pinned in [((40, 10, 10), 24000, 4961, 121),
                                         ((100, 40, 40), 960000, 169781, 1681)]:
        V, T = scenes.kuhn_beam(nx, ny, nz, 0.1, 0.0)
        assert T.shape[0] == nT and V.shape[0] == nV
        fixed = scenes.bc_fixed("hangends", V, 0, 0.4 * 0.1 / (nx * 0.1))
        assert fixed.sum() == pinned


def test_dmat_roundtrip(tmp_path):
    A = np.arange(12, dtype=np.float64).reshape(4, 3) * 0.5
    p = str(tmp_path / "a.dmat")
    meshio.write_dmat(p, A)
    assert np.array_equal(meshio.read_dmat(p), A)


def test_scene_json_files_parse():
    """The JSON scenes shipped as fixtures are valid and use reference keys."""
    import json
    for name in os.listdir(os.path.join(ROOT, "tests", "golden")):
        if name.endswith(".json"):
            d = json.load(open(os.path.join(ROOT, "tests", "golden", name)))
            assert "objects" in d and "material_models" in d


def test_obj_reader_and_remove_unreferenced():
    """square2d.obj: one unreferenced leading vertex, v/vt/vn indices, a quad
    and a face with negative indices. After loading it must be the same
    triangle set as the tri_rectangle generator's 8x3 grid."""
    V, T = meshio.read_obj(os.path.join(ROOT, "tests", "golden", "square2d.obj"))
    assert V.shape == (37, 3) and T.shape == (48, 3)
    assert np.array_equal(V[0], [9, 9, 9])
    assert 0 not in T
    V2, T2 = meshio.remove_unreferenced(V, T)
    assert V2.shape == (36, 3)
    assert np.array_equal(V2, V[1:])  # order kept
    assert np.array_equal(T2, T - 1)
    Vr, Tr = scenes.tri_rectangle(8, 3, 0.125)
    assert np.allclose(V2[:, :2], Vr)
    assert np.allclose(V2[:, 2], 0.25)
    assert sorted(map(tuple, T2.tolist())) == sorted(map(tuple, Tr.tolist()))


def test_scene_loader_transforms():
    """Offset, then the column-major transformation about the centroid; 2D
    drops z and does not rescale; 3D rescales by 1/max first."""
    s = scenes.scene_from_json(os.path.join(ROOT, "tests", "golden", "square2d_scene.json"), dim=2)
    Vr, _ = scenes.tri_rectangle(8, 3, 0.125)
    P = Vr + np.array([1.0, 2.0])
    c = P.mean(axis=0)
    a = np.radians(30.0)
    R = np.array([[np.cos(a), -np.sin(a)], [np.sin(a), np.cos(a)]])
    assert np.allclose(s.V, (P - c) @ R.T + c, atol=1e-14)
    assert s.dim == 2 and np.allclose(s.v0, [0.1, 0.2])
    assert abs(np.linalg.norm(s.V[1] - s.V[0]) - 0.125) < 1e-14  # rigid
    r = scenes.scene_from_json(os.path.join(ROOT, "tests", "golden", "beam_restart_scene.json"))
    V, T = meshio.read_mesh(os.path.join(ROOT, "tests", "golden", "beam6x2x2.mesh"))
    V = V / V.max()
    c3 = V.mean(axis=0)
    t = 0.3
    Rz = np.array([[np.cos(t), -np.sin(t), 0], [np.sin(t), np.cos(t), 0], [0, 0, 1]])
    assert np.allclose(r.V, (V - c3) @ Rz.T + c3, atol=1e-14)
    ids = np.arange(T.shape[0]) % 3 == 0
    assert np.all(r.mat[ids] == scenes.MATERIALS["ARAP"])
    assert np.all(r.mat[~ids] == scenes.MATERIALS["Stable-Neohookean"])
    mu1, _ = scenes.enu_to_lame(np.array([1e6]), 0.4)
    assert np.allclose(r.mu[ids], mu1[0])
    assert r.x0.shape == r.V.shape and r.v0.shape == r.V.shape
    assert np.abs(r.x0 - r.V).max() > 1e-4  # a deformed start


@pytest.mark.parametrize("flip", [False, True])
def test_boundary_faces_close_the_surface(flip):
    """The library's boundary triangles of a Kuhn beam (host code, no GPU):
    2 per boundary cell face, every edge used once in each direction (a
    closed, consistently oriented surface), and the divergence theorem's
    volume of the surface equals the summed tet volumes (outward normals).
    With every tet's corner order flipped the surface must not change
    orientation."""
    from mfem_amd import native
    nx, ny, nz = 4, 3, 2
    V, T = scenes.kuhn_beam(nx, ny, nz, 0.1, 0.1, 7)
    if flip:
        T = T[:, [0, 2, 1, 3]]
    F = native.boundary_faces(V, T)
    assert F.shape[0] == 4 * (nx * ny + ny * nz + nx * nz)
    edges = {}
    for a, b, c in F:
        for e in ((a, b), (b, c), (c, a)):
            edges[e] = edges.get(e, 0) + 1
    assert all(n == 1 for n in edges.values())
    assert all((b, a) in edges for a, b in edges)
    P = V[F]
    vol_surface = np.einsum("ij,ij->i", P[:, 0], np.cross(P[:, 1], P[:, 2])).sum() / 6
    D = V[T[:, 1:]] - V[T[:, [0]]]
    vol_tets = np.abs(np.linalg.det(D)).sum() / 6
    assert abs(vol_surface - vol_tets) < 1e-12 * vol_tets
