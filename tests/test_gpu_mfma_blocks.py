"""Matrix-core (MFMA) variant of the mixed stiffness blocks against the
vector-unit kernel and the FP64 oracle. Needs an MI355X.

Both kernels form vol B^T (Hpsi B) per tet from the same FP32 B and C; the
MFMA variant accumulates the six products of each entry in FP32 on the
matrix cores, so entries agree to FP32 rounding of a six-term sum (relative
to the tet's largest entry; a wrong lane layout would give errors of order
one), the residual f_e is the same code and must be bit-identical, and the
trajectory stays within the mixed-precision bar of the oracle (1e-6
relative 2-norm, the bar of the other mixed-precision tests).
"""
import numpy as np
import pytest

from mfem_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def native():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    from mfem_amd import native as N
    N.lib()
    return N


@pytest.mark.parametrize("material", ["Stable-Neohookean", "ARAP"])
def test_mfma_blocks_match_vector_units(native, oracle, material):
    # 324 tets: the last wave of 64 is partly past the end
    sc = scenes.beam_scene(6, 3, 3, seed=7, mixed=True, E_range=(1e5, 1e6), material=material)
    a = native.Sim(sc, precision="mixed")
    b = native.Sim(sc, precision="mixed")
    b.set_element_kernel(1)
    a.update_system()
    b.update_system()
    ba = a.get("blocks").reshape(-1, 90)
    bb = b.get("blocks").reshape(-1, 90)
    assert ba.shape == (sc.nT, 90)
    scale = np.abs(ba).max(axis=1, keepdims=True)
    assert np.all(scale > 0)
    rel = np.abs(ba - bb) / scale
    assert rel.max() < 1e-5, rel.max()
    assert np.array_equal(a.get("fe"), b.get("fe"))
    o = oracle.Sim(sc)
    for step in range(3):
        b.step()
        o.step()
        xb, xo = b.positions(), o.positions()
        assert np.linalg.norm(xb - xo) / np.linalg.norm(xo) < 1e-6, step


def test_mfma_variant_rejected_without_mixed_stretch(native):
    sc = scenes.beam_scene(3, 2, 2, seed=1, mixed=False)
    s = native.Sim(sc)
    with pytest.raises(native.MfemError, match="mixed-stretch"):
        s.set_element_kernel(1)
    m = native.Sim(scenes.beam_scene(3, 2, 2, seed=1, mixed=True))
    with pytest.raises(native.MfemError, match="0 or 1"):
        m.set_element_kernel(2)
