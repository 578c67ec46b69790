"""C ABI entry points beyond the whole-timestep one, on the GPU: the
linear-solver-level CG on a caller's matrix (mfem_pcg_bsr3) against the
oracle's Eigen-semantics Jacobi CG, and the per-Newton-iteration callback
(the reference's Optimizer::callback, newton_optimizer.cpp:68).

Bars: iteration counts equal (Eigen's stopping rule, one iteration either
way only when |r|^2 lands within rounding of the threshold); the converged
solution within 1e-9 of the oracle's; at tolerance 1e-4 within 2e-6 (the
dot products are summed in a different order); error paths raise before any
CG work."""
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


def csr_to_bsr3(rowptr, colind, vals):
    """Scalar CSR with full 3x3 blocks (the reference's assembled lhs) ->
    (block rowptr, block columns, row-major block values)."""
    n = len(rowptr) - 1
    nb = n // 3
    brp = [0]
    bcol = []
    bval = []
    for i in range(nb):
        rows = [colind[rowptr[3 * i + r]:rowptr[3 * i + r + 1]] for r in range(3)]
        vrow = [vals[rowptr[3 * i + r]:rowptr[3 * i + r + 1]] for r in range(3)]
        cols = rows[0][::3] // 3
        for r in range(3):
            assert np.array_equal(rows[r][::3] // 3, cols)
        for k, c in enumerate(cols):
            blk = np.array([[vrow[r][3 * k + q] for q in range(3)] for r in range(3)])
            assert np.array_equal(rows[0][3 * k:3 * k + 3], 3 * c + np.arange(3))
            bcol.append(c)
            bval.append(blk.reshape(-1))
        brp.append(len(bcol))
    return (np.array(brp, dtype=np.int32), np.array(bcol, dtype=np.int32),
            np.concatenate(bval))


def system_of(native, sc):
    g = native.Sim(sc)
    g.update_system()
    g.solve()  # assembles K and the right-hand side of the first Newton iteration
    rp, ci = g.get_int("rowptr"), g.get_int("colind")
    vals, rhs = g.get("lhs"), g.get("rhs")
    g.close()
    return rp, ci, vals, rhs


def to_dev(*arrs):
    import torch
    return [torch.as_tensor(a).cuda() for a in arrs]


@pytest.mark.parametrize("tol", [1e-12, 1e-4])
def test_pcg_bsr3_matches_oracle(native, oracle, tol):
    sc = scenes.beam_scene(8, 3, 3, seed=21, mixed=True, E_range=(1e5, 1e6))
    rp, ci, vals, rhs = system_of(native, sc)
    brp, bcol, bval = csr_to_bsr3(rp, ci, vals)
    x, it, rr = native.pcg_bsr3(*to_dev(brp, bcol, bval, rhs), tol, 20000)
    x = x.cpu().numpy()
    xo, ito, erro = oracle.jacobi_pcg(rp, ci, vals, rhs, tol, 20000)
    assert abs(it - ito) <= 1, (it, ito)
    bar = 1e-9 if tol < 1e-8 else 2e-6
    assert np.abs(x - xo).max() <= bar * np.abs(xo).max()
    assert rr < tol and erro < tol
    if tol >= 1e-8:  # near 1e-12 both residual recursions sit at rounding level
        assert abs(rr - erro) <= 0.05 * erro


def test_pcg_bsr3_iteration_cap_and_empty(native, oracle):
    sc = scenes.beam_scene(5, 2, 2, seed=22, mixed=False)
    rp, ci, vals, rhs = system_of(native, sc)
    brp, bcol, bval = csr_to_bsr3(rp, ci, vals)
    d = to_dev(brp, bcol, bval, rhs)
    for cap in (0, 1, 7):
        x, it, rr = native.pcg_bsr3(*d, 0.0, cap)
        xo, ito, erro = oracle.jacobi_pcg(rp, ci, vals, rhs, 0.0, cap)
        assert it == ito == cap
        assert np.abs(x.cpu().numpy() - xo).max() <= 1e-12 * max(np.abs(xo).max(), 1e-300)
    import torch
    x, it, rr = native.pcg_bsr3(*to_dev(np.zeros(1, np.int32), np.zeros(0, np.int32),
                                        np.zeros(0), np.zeros(0)), 1e-4, 10)
    assert it == 0 and rr == 0.0 and x.numel() == 0


def test_pcg_bsr3_zero_rhs_and_one_block_row(native, oracle):
    """Eigen's rule for b = 0 (x = 0, no iteration) on a real system, and the
    smallest system, one 3x3 block row, against the oracle and a dense solve."""
    sc = scenes.beam_scene(3, 2, 2, seed=24, mixed=False)
    rp, ci, vals, rhs = system_of(native, sc)
    brp, bcol, bval = csr_to_bsr3(rp, ci, vals)
    x, it, rr = native.pcg_bsr3(*to_dev(brp, bcol, bval, np.zeros_like(rhs)), 1e-4, 100)
    assert it == 0 and rr == 0.0 and not x.cpu().numpy().any()
    A = np.array([[4.0, 1.0, 0.5], [1.0, 3.0, 0.2], [0.5, 0.2, 2.0]])
    b = np.array([1.0, -2.0, 0.5])
    x, it, rr = native.pcg_bsr3(*to_dev(np.array([0, 1], np.int32), np.array([0], np.int32),
                                        A.reshape(-1), b), 1e-12, 100)
    xo, ito, _ = oracle.jacobi_pcg(np.array([0, 3, 6, 9], np.int32),
                                   np.array([0, 1, 2] * 3, np.int32), A.reshape(-1), b, 1e-12, 100)
    assert it == ito
    assert np.abs(x.cpu().numpy() - np.linalg.solve(A, b)).max() <= 1e-12
    assert np.abs(x.cpu().numpy() - xo).max() <= 1e-12


def test_pcg_bsr3_rejects_bad_patterns(native):
    sc = scenes.beam_scene(3, 2, 2, seed=23, mixed=False)
    rp, ci, vals, rhs = system_of(native, sc)
    brp, bcol, bval = csr_to_bsr3(rp, ci, vals)
    bad_col = bcol.copy()
    bad_col[5] = len(brp) + 3
    with pytest.raises(native.MfemError, match="out of range"):
        native.pcg_bsr3(*to_dev(brp, bad_col, bval, rhs), 1e-4, 100)
    bad_rp = brp.copy()
    bad_rp[3], bad_rp[4] = bad_rp[4], bad_rp[3]
    with pytest.raises(native.MfemError, match="row pointer"):
        native.pcg_bsr3(*to_dev(bad_rp, bcol, bval, rhs), 1e-4, 100)
    with pytest.raises(native.MfemError, match="tolerance"):
        native.pcg_bsr3(*to_dev(brp, bcol, bval, rhs), -1.0, 100)


def test_newton_callback_every_iteration(native):
    sc = scenes.beam_scene(6, 3, 3, seed=24, mixed=True, itr_tol=1e-10, max_iters=5000,
                           max_newton=8, newton_tol=1e-9)
    g = native.Sim(sc)
    seen = []
    g.set_callback(lambda it: seen.append((it, g.positions().copy())))
    g.step()
    n = g.stats()["newton_iters"]
    assert [s[0] for s in seen] == list(range(1, n + 1))
    assert np.array_equal(seen[-1][1], g.positions())
    assert not np.array_equal(seen[0][1], sc.V)  # the first iteration moved the beam
    g.set_callback(None)
    seen.clear()
    g.step()
    assert seen == []
