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
