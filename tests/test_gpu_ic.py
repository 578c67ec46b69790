"""GPU parity of "eigen-pcg-IC" (CG with the multicolour block IC(0)
preconditioner) against the FP64 oracle (needs an MI355X).

The solver is checked on the device's own assembled system against the
oracle's scalar restatement of the same algorithm: vertex colours bit-exact,
shift and number of factorization attempts equal (also on stiff beams where
IC(0) breaks down and the shift loop runs), iterates after a fixed number of
iterations to 1e-10, iteration counts at a tolerance equal (within one: the
counts are a threshold test on a rounding-level quantity; 3% on the stiff
beams). Whole timesteps
agree with the oracle to 1e-8 of the largest coordinate in FP64 and to 1e-6
in the mixed FP32/FP64 mode.
"""
import numpy as np
import pytest

from mfem_amd import scenes

pytestmark = pytest.mark.gpu
SOLVER = "eigen-pcg-IC"


@pytest.fixture(scope="module")
def native():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    from mfem_amd import native as N
    N.lib()
    return N


def relerr(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _solved(native, sc, precision="fp64"):
    """Device simulation after one update + solve, with its system in the
    oracle's scalar CSR form."""
    g = native.Sim(sc, precision=precision)
    g.update_system()
    g.solve()
    rp, ci = g.get_int("rowptr"), g.get_int("colind")
    return g, rp, ci, g.get("lhs"), g.get("rhs")


@pytest.mark.parametrize("mixed", [True, False])
def test_colors_shift_and_solution(native, oracle, mixed):
    sc = scenes.beam_scene(4, 3, 3, seed=2, mixed=mixed, solver=SOLVER, itr_tol=1e-10,
                           max_iters=5000)
    print(f"ic solve: mixed={mixed} nT={sc.nT}", flush=True)
    g, rp, ci, K, b = _solved(native, sc)
    L, scale, shift, color, att = oracle.ic_factor_dense(rp, ci, K, 3)
    assert np.array_equal(g.get_int("ic_color"), color)
    assert g.get_int("ic_colors")[0] == color.max() + 1
    assert g.get_int("ic_attempts")[0] == att
    assert g.get("ic_shift")[0] == shift
    x, it, err = oracle.ic_pcg(rp, ci, K, 3, b, 1e-10, 5000)
    it_dev = g.get_int("pcg_iters")[0]
    print(f"  iterations: device {it_dev}, oracle {it}", flush=True)
    assert it > 5 and abs(it_dev - it) <= 1
    assert relerr(g.get("dx"), x) < 1e-8


def test_fixed_iterations_match_oracle(native, oracle):
    sc = scenes.beam_scene(4, 3, 3, seed=5, mixed=True, solver=SOLVER, itr_tol=0.0, max_iters=12)
    g, rp, ci, K, b = _solved(native, sc)
    assert g.get_int("pcg_iters")[0] == 12
    x, it, err = oracle.ic_pcg(rp, ci, K, 3, b, 0.0, 12)
    assert it == 12
    assert relerr(g.get("dx"), x) < 1e-10


@pytest.mark.parametrize("seed", [0, 1])
def test_shift_loop_on_stiff_beam(native, oracle, seed):
    """Per-tet E in [1e4, 1e8], nu 0.49, ARAP/SNH: IC(0) breaks down and the
    shift doubles until it does not. The iteration count at 1e-8 agrees to
    3% here (measured 91 vs 89 on seed 1): the ill-conditioned system
    amplifies the rounding difference between the device's block elimination
    and the oracle's scalar one."""
    sc = scenes.config_c5(6, 3, 3, seed=seed, itr_tol=1e-8, max_iters=5000)
    sc.solver = SOLVER
    print(f"ic stiff beam: seed {seed}", flush=True)
    g, rp, ci, K, b = _solved(native, sc)
    L, scale, shift, color, att = oracle.ic_factor_dense(rp, ci, K, 3)
    assert att > 1
    assert g.get_int("ic_attempts")[0] == att
    assert g.get("ic_shift")[0] == shift
    x, it, err = oracle.ic_pcg(rp, ci, K, 3, b, 1e-8, 5000)
    assert abs(g.get_int("pcg_iters")[0] - it) <= max(1, 0.03 * it)
    assert relerr(g.get("dx"), x) < 1e-6


@pytest.mark.parametrize("mixed,precision,tol", [(True, "fp64", 1e-8), (False, "fp64", 1e-8),
                                                 (True, "mixed", 1e-6)])
def test_timesteps_match_oracle(native, oracle, mixed, precision, tol):
    sc = scenes.beam_scene(4, 2, 2, seed=1, mixed=mixed, solver=SOLVER, itr_tol=1e-12,
                           max_iters=5000, newton_tol=1e-9)
    g = native.Sim(sc, precision=precision)
    o = oracle.Sim(sc)
    for k in range(3):
        print(f"ic timestep {k} mixed={mixed} {precision}", flush=True)
        g.step()
        o.step()
        assert relerr(g.positions(), o.positions()) < tol, k
        if precision == "fp64":
            assert g.stats()["newton_iters"] == o.stats()["newton_iters"]


def test_2d_mixed_timesteps(native, oracle):
    sc = scenes.rect2d_scene(16, 5, 0.1, E=1e5, solver=SOLVER, itr_tol=1e-12, max_iters=20000,
                             max_newton=20, newton_tol=1e-9)
    sc.mixed = True
    g = native.Sim(sc, precision="fp64")
    o = oracle.Sim(sc)
    for k in range(3):
        print(f"ic 2d timestep {k}", flush=True)
        g.step()
        o.step()
        assert relerr(g.positions(), o.positions()) < 1e-8, k
    assert g.get_int("ic_colors")[0] > 1


def test_matches_jacobi_cg(native):
    """Both device CG solvers, solved tightly, give the same timesteps, the IC
    one in fewer iterations."""
    kw = dict(seed=3, mixed=True, itr_tol=1e-12, max_iters=20000, newton_tol=1e-9)
    a = native.Sim(scenes.beam_scene(6, 3, 3, solver=SOLVER, **kw), precision="fp64")
    j = native.Sim(scenes.beam_scene(6, 3, 3, solver="eigen-pcg-diag", **kw), precision="fp64")
    ia = ij = 0
    for _ in range(2):
        a.step()
        j.step()
        ia += a.stats()["pcg_iters"]
        ij += j.stats()["pcg_iters"]
    disp = np.abs(j.positions() - j.scene.V).max()
    assert np.abs(a.positions() - j.positions()).max() < 1e-7 * disp
    assert ia < ij
