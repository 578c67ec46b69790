"""The benchmark configurations at their full sizes on the GPU, against the
FP64 oracle's committed trajectories (tests/golden/config_*.npz, written by
tests/golden/make_config_fixtures.py) and against the device's own FP64
mode.

Rounding sensitivity. At the reference's default solver settings (5 Newton
iterations, CG to 1e-4) the 960k-tet beam's trajectory becomes sensitive to
rounding once the material projection starts to act (from about timestep
12): two FP64 runs that differ only in the summation order of the CG dot
products (two CG kernel shapes) drift apart by a factor of about 3 per
timestep, to ~2e-6 (relative 2-norm of the positions) by timestep 24. No
implementation can therefore be held to 1e-6 of another over all 25
timesteps. The bar used is: within 1e-6, or within 3x that FP64-vs-FP64
spread, measured in the same test on the same vertices.
"""
import os

import numpy as np
import pytest

from mfem_amd import scenes
from tests.golden import make_config_fixtures as F

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def native():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    from mfem_amd import native as N
    N.lib()
    return N


def rel2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


CG_SHAPES = [(4, 4, 256), (4, 4, 512), (8, 2, 256), (8, 2, 1024), (4, 2, 256), (4, 2, 1024),
             (16, 1, 1024), (8, 1, 256)]


def other_shapes(sim, n=1):
    """n CG kernel shapes other than the one the autotuner picked for `sim`
    (other lanes per row and workgroup sizes: other summation orders)."""
    _, _, shape = sim.cg_tuning()
    return [v for v in (3, 6, 0, 2) if CG_SHAPES[v] != tuple(shape[:3])][:n]


def rounding_floor_failures(sim, sc):
    """Line searches of the last step that ran out of halvings where the
    Newton iteration had reached FP64 rounding: energy change below 1e-10
    relative and max |dx|, |ds| below 1e-6. There E(alpha) < E(0) compares
    two sums that differ by their own rounding, and whether some halving
    passes is chance in either precision (the FP64 run takes 0-4 halvings at
    such iterations); the skipped step moves no position by more than 1e-6."""
    h = sim.newton_history()
    cols = sim.HISTORY_COLUMNS
    hv, res, dec = (cols.index(c) for c in ("ls_halvings", "energy_res", "newton_dec"))
    return int(np.sum((h[:, hv] >= sc.ls_iters) & (h[:, res] < 1e-10) & (h[:, dec] < 1e-6)))


def test_c3_mixed_and_fp64_25_steps(native):
    """The precision config (100x40x40x6 = 960,000 tets) for the 25 timesteps
    the benchmark runs: mixed vs FP64 device, and the FP64 device vs the
    oracle's trajectory, each to max(1e-6, 3 x the FP64 rounding spread); no
    line search fails in mixed mode that does not fail in FP64, except at
    iterations that have reached rounding (rounding_floor_failures); Newton
    counts equal. The spread is the largest of the three pairs of FP64 runs
    with three CG kernel shapes (one pair is a single sample of the
    trajectory's sensitivity to summation order)."""
    data = np.load(os.path.join(GOLDEN, "config_c3.npz"))
    sc = scenes.config_c3()
    ids, hot = data["sub_ids"], data["hot_ids"]
    gm = native.Sim(sc, precision="mixed")
    ga = native.Sim(sc, precision="fp64")
    gb = native.Sim(sc, precision="fp64")
    gc = native.Sim(sc, precision="fp64")
    worst = []
    for k in range(int(data["steps"])):
        gm.step()
        ga.step()
        if k == 0:
            vb, vc = other_shapes(ga, 2)
            gb.set_cg_shape(vb, 0)
            gc.set_cg_shape(vc, 0)
        gb.step()
        gc.step()
        xm, xa, xb, xc = gm.positions(), ga.positions(), gb.positions(), gc.positions()
        sm, sa = gm.stats(), ga.stats()
        spread = max(rel2(xb, xa), rel2(xc, xa), rel2(xc, xb))
        err = rel2(xm, xa)
        assert err <= max(1e-6, 3 * spread), (k, err, spread)
        assert sm["ls_fail"] - sa["ls_fail"] <= rounding_floor_failures(gm, sc), \
            (k, sm["ls_fail"], sa["ls_fail"], gm.newton_history())
        assert sm["newton_iters"] == sa["newton_iters"] == data["newton"][k], k
        # FP64 device vs the oracle on the seeded subset (every step) and the
        # vertices next to the clamp (every 5th step)
        xo = data["sub_x"][k]
        def spread_on(sel):
            return max(rel2(xb[sel], xa[sel]), rel2(xc[sel], xa[sel]), rel2(xc[sel], xb[sel]))
        e_sub = rel2(xa[ids], xo)
        assert e_sub <= max(1e-8, 3 * spread_on(ids)), (k, e_sub)
        e_hot = None
        if k in data["hot_steps"]:
            j = list(data["hot_steps"]).index(k)
            e_hot = rel2(xa[hot], data["hot_x"][j])
            assert e_hot <= max(1e-8, 3 * spread_on(hot)), (k, e_hot)
            print(f"c3 step {k}: fp64 vs oracle next to the clamp {e_hot:.1e}, fp64 spread "
                  f"there {spread_on(hot):.1e}")
        worst.append((k, err, spread, e_sub))
    # where the FP64 spread is small, the plain 1e-6 bar holds
    for k, err, spread, _ in worst:
        if spread < 1e-7:
            assert err < 1e-6, (k, err)
    print("c3 (step, mixed-fp64, fp64 spread, fp64-oracle subset):",
          [(k, f"{e:.1e}", f"{s:.1e}", f"{o:.1e}") for k, e, s, o in worst])


def test_c2_full_size_matches_oracle(native):
    """The plain FP64 config as configured (40x10x10x6 = 24,000 tets, Jacobi
    CG to rel-res 1e-10, 10 timesteps) against the oracle's trajectory."""
    data = np.load(os.path.join(GOLDEN, "config_c2.npz"))
    sc, steps = F.config_scene("c2")
    ids = data["sub_ids"]
    g = native.Sim(sc)
    for k in range(steps):
        g.step()
        x = g.positions()
        assert g.stats()["newton_iters"] == data["newton"][k], k
        assert np.abs(x[ids] - data["sub_x"][k]).max() < 1e-8 * np.abs(data["sub_x"][k]).max(), k
        assert abs(np.linalg.norm(x) - data["xnorm"][k]) < 1e-10 * data["xnorm"][k], k
        if k in data["hot_steps"]:
            j = list(data["hot_steps"]).index(k)
            ref = data["hot_x"][j]
            assert np.abs(x[data["hot_ids"]] - ref).max() < 1e-8 * np.abs(ref).max(), k


def test_c5_full_size_properties(native):
    """The stiff config at full size (130x50x50x6 = 1.95M tets, ARAP / SNH, E in
    [1e4, 1e8], nu 0.49), where CG stops at its 500-iteration cap: Newton
    counts equal in both precisions, E(0) decreases over every step's Newton
    iterations, the cap is respected and the residual reported, and mixed
    stays within 1e-6 of FP64 over 3 timesteps."""
    sc = scenes.config_c5()
    gm = native.Sim(sc, precision="mixed")
    g64 = native.Sim(sc, precision="fp64")
    for k in range(3):
        gm.step()
        g64.step()
        for g in (gm, g64):
            st = g.stats()
            hist = g.newton_history()
            assert np.all(np.diff(hist[:, 1]) < 0), (k, hist[:, 1])
            assert st["pcg_iters"] <= sc.max_iters * st["newton_iters"]
            assert np.isfinite(st["rel_residual"]) and st["rel_residual"] > 0
            assert st["ls_fail"] == 0
        assert gm.stats()["newton_iters"] == g64.stats()["newton_iters"], k
        err = rel2(gm.positions(), g64.positions())
        assert err < 1e-6, (k, err)


def test_c4_full_batch(native, oracle):
    """The batching config as configured (256 beams, 2e3-6e4 tets each, mixed,
    10 timesteps, one batch): beams spread over the size range equal their
    single-simulation device runs (the batch keeps per-beam CG scalars, line
    search and termination; only the order of its partial sums differs), and
    the 4 smallest beams follow the FP64 oracle to 1e-6."""
    scs = scenes.config_c4()
    nT = np.array([s.nT for s in scs])
    order = np.argsort(nT)
    pick = [int(k) for k in order[[0, 1, 2, 3, 64, 128, 192, 255]]]
    b = native.Batch(scs, precision="mixed")
    singles = {k: native.Sim(scs[k], precision="mixed") for k in pick}
    orc = {k: oracle.Sim(scs[k]) for k in pick[:4]}
    for step in range(10):
        b.step()
        st = b.stats()
        for k, s in singles.items():
            s.step()
            assert rel2(b.positions(k), s.positions()) < 1e-8, (step, k)
            assert st["newton_iters"][k] == s.stats()["newton_iters"], (step, k)
        for k, o in orc.items():
            o.step()
            assert rel2(b.positions(k), o.positions()) < 1e-6, (step, k)


def _clamped_tets(oracle, sc, s):
    """Tets whose unprojected SNH Hessian fails the projection's positive-
    definiteness test (Cholesky of H - 1e-6 I), i.e. where the projection acts."""
    n = 0
    for e in range(sc.nT):
        _, _, H = oracle.material_s6(0, sc.mu[e], sc.la[e], s[e], False)
        try:
            np.linalg.cholesky(H - 1e-6 * np.eye(6))
        except np.linalg.LinAlgError:
            n += 1
    return n


def test_mixed_with_projection_active(native, oracle):
    """A beam pulled apart by its end groups (the reference's stretch script,
    velocity -1) at the reference's default solver settings: from timestep 4
    on, tens of tets have an indefinite SNH Hessian, so the PSD projection
    and the inverse of the projected matrix carry a spread of eigenvalues of
    ~1e12. Mixed mode must follow the oracle there: no extra line-search
    failures, Newton decrements within 5%, positions within 1e-4 (Newton stops
    at 5 iterations with decrements of ~1e-2, so the CG tolerance 1e-4 shows
    up at ~1e-5); the FP64 device within 1e-6."""
    sc = scenes.beam_scene(8, 3, 3, seed=5, mixed=True, E_range=(1e5, 1e6))
    sc.bc = dict(type="stretch", axis=0, ratio=0.05, velocity=-1.0)
    sc.fixed = scenes.bc_fixed("stretch", sc.V, 0, 0.05)
    gm = native.Sim(sc, precision="mixed")
    g64 = native.Sim(sc)
    o = oracle.Sim(sc)
    clamped = []
    for k in range(12):
        gm.step()
        g64.step()
        o.step()
        xo = o.positions()
        so, sm = o.stats(), gm.stats()
        clamped.append(_clamped_tets(oracle, sc, o.get("s").reshape(-1, 6)))
        assert sm["ls_fail"] <= so["ls_fail"], k
        if so["last_decrement"] > 1e-6:  # below, the decrement is the Newton loop's noise
            assert abs(sm["decrement"] - so["last_decrement"]) <= 0.05 * so["last_decrement"], k
        assert rel2(gm.positions(), xo) < 1e-4, k
        assert rel2(g64.positions(), xo) < 1e-6, k
    assert min(clamped[6:]) > 10, clamped
