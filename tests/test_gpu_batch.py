"""Batched independent simulations (mfem_batch_*) against each simulation run
alone: the FP64 oracle and the single-simulation device path. Needs an MI355X.

Each simulation in a batch must follow the single-simulation algorithm (its
own CG scalars and Eigen stopping rule, line search and Newton termination),
so positions agree with the oracle to the same bar as a lone simulation
(1e-8 relative, FP64) and Newton counts are equal.
"""
import numpy as np
import pytest

from mfem_amd import scenes

pytestmark = pytest.mark.gpu
POS_TOL = 1e-8


@pytest.fixture(scope="module")
def native():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    from mfem_amd import native as N
    N.lib()
    return N


def relerr(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def small_batch(**kw):
    """Five beams of different sizes, stiffness and material."""
    sizes = [(4, 2, 2), (6, 3, 3), (5, 2, 3), (8, 3, 2), (3, 2, 2)]
    E = [1e5, 3e5, 1e6, 2e5, 5e5]
    mats = ["Stable-Neohookean", "Stable-Neohookean", "ARAP", "Stable-Neohookean", "Corot"]
    out = []
    for k, (nx, ny, nz) in enumerate(sizes):
        out.append(scenes.beam_scene(nx, ny, nz, seed=20 + k, mixed=True, E=E[k], material=mats[k],
                                     **kw))
    return out


def test_batch_matches_oracle_fp64(native, oracle):
    kw = dict(itr_tol=1e-12, max_iters=20000, max_newton=20, newton_tol=1e-9)
    scs = small_batch(**kw)
    batch = native.Batch(scs, precision="fp64")
    orc = [oracle.Sim(sc) for sc in scs]
    for step in range(3):
        batch.step()
        st = batch.stats()
        for k, o in enumerate(orc):
            o.step()
            assert relerr(batch.positions(k), o.positions()) < POS_TOL, (step, k)
            assert st["newton_iters"][k] == o.stats()["newton_iters"], (step, k)


def test_batch_default_tolerances_mixed(native, oracle):
    """Reference defaults (CG tol 1e-4 / 500, 5 Newton iterations), mixed
    precision: every beam within 1e-6 (relative 2-norm) of its oracle run."""
    scs = small_batch()
    batch = native.Batch(scs, precision="mixed")
    orc = [oracle.Sim(sc) for sc in scs]
    for step in range(3):
        batch.step()
        for k, o in enumerate(orc):
            o.step()
            xo = o.positions()
            assert np.linalg.norm(batch.positions(k) - xo) / np.linalg.norm(xo) < 1e-6, (step, k)


@pytest.mark.parametrize("variant", [0, 3, 6])
def test_batch_equals_single_simulations(native, variant):
    """FP64 batch vs each beam as its own device simulation, several CG
    shapes and rows per workgroup (workgroups of a beam split its rows)."""
    kw = dict(itr_tol=1e-12, max_iters=20000, max_newton=20, newton_tol=1e-9)
    scs = small_batch(**kw)
    batch = native.Batch(scs, precision="fp64")
    batch.set_cg_shape(variant, 16)
    singles = [native.Sim(sc) for sc in scs]
    for step in range(2):
        batch.step()
        for k, s in enumerate(singles):
            s.step()
            assert relerr(batch.positions(k), s.positions()) < POS_TOL, (step, k)


def test_batch_rejects_mismatched_settings(native):
    a = scenes.beam_scene(3, 2, 2, seed=1, mixed=True)
    b = scenes.beam_scene(3, 2, 2, seed=2, mixed=True, dt=0.01)
    with pytest.raises(native.MfemError, match="must be the same"):
        native.Batch([a, b])
    c = scenes.beam_scene(3, 2, 2, seed=3, mixed=False)
    with pytest.raises(native.MfemError, match="only mixed-stretch"):
        native.Batch([a, c])


def test_batch_graph_replay_is_bit_identical_at_256_beams(native):
    """The batching config as configured (256 beams, mixed precision): the
    hipGraph replay of the update + assembly phase and of CG iteration pairs
    gives bit-identical positions and counts to launching every kernel, over
    several steps (each replay follows eager work and synchronisations), and
    submits one update graph per Newton iteration instead of four launches
    and one CG graph per 16 iterations."""
    scs = scenes.config_c4()
    g = native.Batch(scs, precision="mixed")
    e = native.Batch(scs, precision="mixed")
    g.set_graphs(True)
    e.set_graphs(False)
    for b in (g, e):  # one CG shape for both (the autotuner could pick two)
        b.set_cg_shape(2, 128)
    for step in range(4):
        g.step()
        e.step()
        sg, se = g.stats(), e.stats()
        for key in ("newton_iters", "pcg_iters", "refinements", "ls_fail"):
            assert np.array_equal(sg[key], se[key]), (step, key)
        for k in range(g.count):
            assert np.array_equal(g.positions(k), e.positions(k)), (step, k)
    ug, cg, ng = g.launch_counts()
    ue, ce, ne = e.launch_counts()
    assert ng == ne > 0
    assert ue == 4 * ne
    # the first step runs eagerly (its buffers are allocated), then one graph
    assert ug < ue and ug >= ng
    assert cg < 0.2 * ce
