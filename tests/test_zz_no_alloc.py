"""mfem_pcg_bsr3_noalloc (the CG entry point that allocates no device memory:
its work arrays are a fixed device array of the library) against
mfem_pcg_bsr3 on the same systems: the same kernels on the same launch shape,
so the results must be bit-identical. Two calls in a row, the second on a
smaller system (the fixed array keeps the first call's contents), and a
system above the row limit is rejected before any device work."""
import ctypes as C

import numpy as np
import pytest

from mfem_amd import scenes
from tests.test_gpu_abi import csr_to_bsr3, system_of, to_dev

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def native():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    from mfem_amd import native as N
    N.lib()
    return N


def _bsr_system(native, sc):
    rp, ci, vals, rhs = system_of(native, sc)
    return to_dev(*csr_to_bsr3(rp, ci, vals), rhs)


def _solve_both(native, dev, tol, cap):
    import torch
    n = dev[3].numel()
    xa = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
    xn = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
    _, ita, rra = native.pcg_bsr3(*dev, tol, cap, x=xa)
    _, itn, rrn = native.pcg_bsr3(*dev, tol, cap, noalloc=True, x=xn)
    torch.cuda.synchronize()
    return xa.cpu().numpy(), ita, rra, xn.cpu().numpy(), itn, rrn


def test_noalloc_equals_allocating_entry_point(native):
    big = _bsr_system(native, scenes.beam_scene(10, 4, 4, seed=31, mixed=True,
                                                E_range=(1e5, 1e6)))
    small = _bsr_system(native, scenes.beam_scene(4, 2, 2, seed=32, mixed=False))
    for dev, tol, cap in ((big, 1e-10, 20000), (small, 1e-4, 500)):
        xa, ita, rra, xn, itn, rrn = _solve_both(native, dev, tol, cap)
        assert np.isfinite(xn).all()
        assert itn == ita and itn > 0
        assert rrn == rra and rrn < tol
        assert np.array_equal(xn, xa)


def test_noalloc_rejects_systems_above_the_limit(native):
    nb = native.PCG_NOALLOC_MAX_ROWS // 3 + 1
    it, rr = C.c_int(-1), C.c_double(-1.0)
    rc = native.lib().mfem_pcg_bsr3_noalloc(nb, 0, None, None, None, None, None, 1e-4, 10,
                                            C.byref(it), C.byref(rr), None)
    assert rc != 0
    assert "no-allocation limit" in native.lib().mfem_last_error().decode()
    assert it.value == -1
