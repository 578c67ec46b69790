"""Oracle trajectories of the benchmark configurations at their full sizes.

Run:  python -m tests.golden.make_config_fixtures [c2] [c3] [--threads N]

Writes tests/golden/config_<name>.npz with, after every timestep of the FP64
oracle:
  * the positions of a seeded subset of vertices (`sub_x`, steps x k x 3) and
    the subset's vertex ids (`sub_ids`),
  * the 2-norm of all positions (`xnorm`) and of the displacement from the
    rest shape (`unorm`),
  * Newton iterations, CG iterations, line-search failures, the last Newton
    decrement and the last energy of the step;
  * every 5th step, the positions of all free vertices within 0.25 (2.5
    cells) of the clamped face (`hot_x`, `hot_ids`, `hot_steps`): the
    strongest strains, where the material projection acts first.

The scenes are the ones the benchmark runs (mfem_amd.scenes.config_c2 and
config_c3: seeded Kuhn beams, splitmix64 streams), so the fixture is
reproducible. Cost on 8 host cores: C2 about 10 s, C3 (25 timesteps of the
960,000-tet beam) about 15 minutes.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from mfem_amd import scenes  # noqa: E402

SUBSET = 2048          # vertices whose positions are stored every step
SUBSET_STREAM = 41     # splitmix64 stream of the subset draw


def config_scene(name):
    """(scene, timesteps) of a benchmark configuration."""
    if name == "c2":
        return scenes.config_c2(), 10
    if name == "c3":
        return scenes.config_c3(), 25
    raise ValueError(name)


def subset_ids(nV, seed=0, k=SUBSET):
    """k distinct vertex ids drawn from a seeded stream (sorted)."""
    u = scenes.splitmix64_uniform(seed, SUBSET_STREAM, 4 * k)
    ids = np.unique((u * nV).astype(np.int64))[:k]
    return np.sort(ids)


def hot_ids(sc):
    """Free vertices within 0.25 of the clamped face x = max (sorted)."""
    x = sc.V[:, 0]
    return np.nonzero((x >= x.max() - 0.25) & (sc.fixed == 0))[0]


HOT_EVERY = 5


def run(name, threads=0, on_step=None):
    from oracle import oracle as O
    sc, steps = config_scene(name)
    ids = subset_ids(sc.nV)
    hot = hot_ids(sc)
    sim = O.Sim(sc, threads=threads)
    rec = {k: [] for k in ("sub_x", "xnorm", "unorm", "newton", "pcg", "ls_fail",
                           "decrement", "energy", "hot_x", "hot_steps")}
    for k in range(steps):
        t0 = time.perf_counter()
        sim.step()
        st = sim.stats()
        x = sim.positions()
        rec["sub_x"].append(x[ids])
        rec["xnorm"].append(np.linalg.norm(x))
        rec["unorm"].append(np.linalg.norm(x - sc.V))
        rec["newton"].append(st["newton_iters"])
        rec["pcg"].append(st["pcg_iters_total"])
        rec["ls_fail"].append(st["ls_fail"])
        rec["decrement"].append(st["last_decrement"])
        rec["energy"].append(st["last_energy"])
        if k % HOT_EVERY == HOT_EVERY - 1:
            rec["hot_x"].append(x[hot])
            rec["hot_steps"].append(k)
        print(f"{name} step {k}: {time.perf_counter() - t0:.1f} s newton {st['newton_iters']} "
              f"cg {st['pcg_iters_total']} ls_fail {st['ls_fail']} dec {st['last_decrement']:.3e}",
              flush=True)
        if on_step is not None:
            on_step(k, sim, sc)
    out = {k: np.asarray(v) for k, v in rec.items()}
    out["sub_ids"] = ids
    out["hot_ids"] = hot
    out["steps"] = np.int64(steps)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("configs", nargs="*", default=["c2", "c3"])
    ap.add_argument("--threads", type=int, default=0)
    args = ap.parse_args()
    for name in args.configs:
        out = run(name, args.threads)
        path = os.path.join(HERE, f"config_{name}.npz")
        np.savez_compressed(path, **out)
        print("wrote", path, flush=True)


if __name__ == "__main__":
    main()
