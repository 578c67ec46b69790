"""Regenerate tests/golden/trajectories.npz from the FP64 oracle.

Run:  python -m tests.golden.make_golden
Each case is (scene, number of timesteps); the fixture holds the vertex
positions after every timestep. The scenes are seeded (mfem_amd.scenes uses a
counter-based splitmix64 stream), so the fixture is reproducible.
"""
import os

import numpy as np

from mfem_amd import scenes


def _c1():
    return scenes.config_c1(), 20


def _beam_mixed():
    return scenes.beam_scene(4, 2, 2, seed=11, mixed=True, solver="eigen-llt",
                             max_newton=10, newton_tol=1e-9), 5


def _beam_plain():
    return scenes.beam_scene(4, 2, 2, seed=11, mixed=False, solver="eigen-llt",
                             max_newton=10, newton_tol=1e-9), 5


def _beam_mixed_pcg_percell():
    return scenes.beam_scene(5, 2, 2, seed=12, E_range=(1e5, 1e6), mixed=True,
                             itr_tol=1e-10, max_iters=5000), 4


CASES = {
    "c1": _c1,
    "beam_mixed": _beam_mixed,
    "beam_plain": _beam_plain,
    "beam_mixed_pcg": _beam_mixed_pcg_percell,
}


def main():
    from oracle import oracle as O
    out = {}
    for name, fn in CASES.items():
        sc, steps = fn()
        sim = O.Sim(sc)
        for k in range(steps):
            sim.step()
            out[f"{name}_x{k}"] = sim.positions()
        print(name, sim.stats())
    path = os.path.join(os.path.dirname(__file__), "trajectories.npz")
    np.savez_compressed(path, **out)
    print("wrote", path)


if __name__ == "__main__":
    main()
