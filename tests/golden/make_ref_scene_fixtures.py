"""Oracle trajectories of two of the reference's own scene files, loaded
unchanged (tests/golden/ref_scenes/scenes/3D/*.json with the meshes and state
files they name, copied from the reference's scenes/ and models/ folders).

Run:  python -m tests.golden.make_ref_scene_fixtures [--threads N]

Writes tests/golden/ref_scene_<name>.npz with the positions after every
timestep (`x`, steps x nV x 3), the Newton iterations, line-search failures,
last Newton decrement and energy residual of every step.

  * beam_twist (751 vertices, 2,998 tets, mixed-stretch, "cholmod", the
    twist script, duplicate "path0" keys): 5 timesteps.
  * gummy_bear_deformed (5,701 vertices, 24,393 tets, displacement-only
    "stretch", "cholmod", hangends, per-element materials with E = 6e4 and
    1e14, initial_state restart from sim_x0_0013 / sim_v_0013): 3 timesteps.
    Newton runs to its 100-iteration cap in every step (the 1e14 region
    makes the system very stiff); about 5 minutes per step on 8 cores.
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

SCENE_DIR = os.path.join(HERE, "ref_scenes", "scenes", "3D")
CASES = {"beam_twist": 5, "gummy_bear_deformed": 3}


def scene_path(name):
    return os.path.join(SCENE_DIR, name + ".json")


def run(name, threads=0):
    from oracle import oracle as O
    sc = scenes.scene_from_json(scene_path(name))
    sim = O.Sim(sc, threads=threads)
    rec = {k: [] for k in ("x", "newton", "ls_fail", "decrement", "res")}
    for k in range(CASES[name]):
        t0 = time.perf_counter()
        sim.step()
        st = sim.stats()
        rec["x"].append(sim.positions())
        rec["newton"].append(st["newton_iters"])
        rec["ls_fail"].append(st["ls_fail"])
        rec["decrement"].append(st["last_decrement"])
        rec["res"].append(st["last_res"])
        print(f"{name} step {k}: {time.perf_counter() - t0:.1f} s, newton {st['newton_iters']}, "
              f"decrement {st['last_decrement']:.3e}", flush=True)
    return {k: np.asarray(v) for k, v in rec.items()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("names", nargs="*", default=list(CASES))
    ap.add_argument("--threads", type=int, default=0)
    args = ap.parse_args()
    for name in args.names:
        out = run(name, args.threads)
        path = os.path.join(HERE, f"ref_scene_{name}.npz")
        np.savez_compressed(path, **out)
        print("wrote", path, flush=True)


if __name__ == "__main__":
    main()
