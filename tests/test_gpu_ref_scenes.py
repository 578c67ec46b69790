"""Two of the reference's own scene files, run unchanged through the device's
JSON loader (mfem_sim_create_json) against the FP64 oracle's trajectories
(tests/golden/ref_scene_*.npz from tests/golden/make_ref_scene_fixtures.py):

  * scenes/3D/beam_twist.json with models/mesh/beam.mesh: mixed-stretch,
    "cholmod", the twist script, duplicate "path0" keys, a top-level
    "boundary_condition" string the loader ignores;
  * scenes/3D/gummy_bear_deformed.json: displacement-only "stretch",
    "cholmod", hangends, per-element materials from a DMAT with E = 6e4 and
    1e14, and an initial_state restart (x and v DMATs).

"cholmod" runs as the device's dense FP64 Cholesky, the oracle's as a sparse
Cholesky; both are exact factorizations, so only rounding differs. The
mixed FP32/FP64 mode is selected by the loader's "precision" key on a copy
of the scene; with a direct solver it stores everything in FP64 (FP32
storage only pays in the iterative solvers, and an FP32-rounded K of the
gummy bear, E = 6e4 next to 1e14, is not positive definite), so its
trajectory equals the FP64 one bit for bit."""
import json
import os
import shutil

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
SCENES = os.path.join(GOLDEN, "ref_scenes")


@pytest.fixture(scope="module")
def native():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    from mfem_amd import native as N
    N.lib()
    return N


def rel2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def mixed_copy(tmp_path, name):
    """The scene tree copied under tmp_path with "precision": "mixed" added."""
    dst = tmp_path / "ref_scenes"
    shutil.copytree(SCENES, dst)
    path = dst / "scenes" / "3D" / f"{name}.json"
    d = json.loads(path.read_text())
    d["precision"] = "mixed"
    path.write_text(json.dumps(d))
    return str(path)


def scene_json(name):
    return os.path.join(SCENES, "scenes", "3D", f"{name}.json")


def test_beam_twist_scene_unchanged(native, tmp_path):
    data = np.load(os.path.join(GOLDEN, "ref_scene_beam_twist.npz"))
    g = native.Sim(json_path=scene_json("beam_twist"))
    gm = native.Sim(json_path=mixed_copy(tmp_path, "beam_twist"))
    for k in range(len(data["newton"])):
        g.step()
        gm.step()
        xo = data["x"][k]
        assert g.stats()["newton_iters"] == data["newton"][k], k
        assert np.abs(g.positions() - xo).max() < 1e-8 * np.abs(xo).max(), k
        assert np.array_equal(gm.positions(), g.positions()), k
    # the twist script moved the ends
    assert np.abs(data["x"][-1] - data["x"][0]).max() > 1e-3


def test_gummy_bear_scene_unchanged(native, tmp_path):
    """Newton runs to its 100-iteration cap in every step on both sides (the
    1e14 region) and stops with Newton decrements (max |dx|) of 2.5e-5 to
    2.2e-4, so the positions carry an unresolved Newton update of that size,
    which the two sides' rounding (dense vs sparse Cholesky of a matrix whose
    stiffnesses span ten decades) shifts, and which the next step inherits.
    The bar: every vertex within the sum of the final Newton decrements of the
    steps so far, and the positions within 1e-4 (relative 2-norm)."""
    data = np.load(os.path.join(GOLDEN, "ref_scene_gummy_bear_deformed.npz"))
    g = native.Sim(json_path=scene_json("gummy_bear_deformed"))
    gm = native.Sim(json_path=mixed_copy(tmp_path, "gummy_bear_deformed"))
    errs = []
    dec_sum = 0.0
    for k in range(len(data["newton"])):
        g.step()
        gm.step()
        xo = data["x"][k]
        st = g.stats()
        assert st["newton_iters"] == data["newton"][k], k
        assert st["ls_fail"] <= data["ls_fail"][k] + 1, k
        e64 = rel2(g.positions(), xo)
        dmax = float(np.abs(g.positions() - xo).max())
        dec_sum += float(data["decrement"][k])
        errs.append((k, e64, dmax, dec_sum))
        assert dmax <= dec_sum, (k, dmax, dec_sum)
        assert e64 < 1e-4, (k, e64)
        assert np.array_equal(gm.positions(), g.positions()), k
    print("gummy bear (step, fp64 vs oracle 2-norm, max abs, summed oracle decrements):", errs)
