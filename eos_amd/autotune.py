"""Autotuner for the traversal launch configuration.

Searches workgroup size, per-lane LDS stack depth, persistent waves per CU,
the refill threshold (how many idle lanes a wave refills at once) and the ray
batch (how many items a wave claims per fetch) on a given workload, and
persists the winner as JSON.  ``load`` returns a saved configuration for
``Scene.set_launch_config``.
"""
from __future__ import annotations

import itertools
import json
import os
import time

import numpy as np

from . import _lib as E

DEFAULT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_launch.json")

GRID = {
    "workgroup": (128, 256),
    "stack_lds": (8, 16),
    "waves_per_cu": (16, 20, 32),
    "refill": (16,),
    "batch": (64, 128),
    "primary_packets": (0, 1),
}


def _time_batch(scenes, depth, reps, use_graph):
    E.render_batch(scenes, depth=depth, use_graph=use_graph)        # warm-up / capture
    ms = []
    for _ in range(reps):
        _, st = E.render_batch(scenes, depth=depth, use_graph=use_graph)
        ms.append(st["render_ms"])
    return float(np.median(ms))


def _time_single(scene, depth, reps):
    scene.render(depth=depth, want_rgba=False)
    return float(np.median([scene.render(depth=depth, want_rgba=False)["stats"]["render_ms"] for _ in range(reps)]))


def tune(scenes, depth: int, reps: int = 3, grid=None, batched: bool = True, use_graph: bool = True,
         path: str | None = DEFAULT_PATH, workload: str = "") -> dict:
    """Times every configuration of ``grid`` and returns (and saves) the fastest.

    batched=True times eos_render_batch over ``scenes`` (the batch uses scenes[0]'s
    configuration); batched=False times eos_render of scenes[0]."""
    grid = grid or GRID
    keys = list(grid)
    results = []
    t0 = time.perf_counter()
    for vals in itertools.product(*(grid[k] for k in keys)):
        cfg = dict(zip(keys, vals))
        scenes[0].set_launch_config(**cfg)
        ms = _time_batch(scenes, depth, reps, use_graph) if batched else _time_single(scenes[0], depth, reps)
        results.append((ms, cfg))
    results.sort(key=lambda r: r[0])
    best_ms, best = results[0]
    scenes[0].set_launch_config(**best)
    out = {"config": best, "ms": best_ms, "worst_ms": results[-1][0], "tried": len(results),
           "seconds": time.perf_counter() - t0, "workload": workload, "batched": batched,
           "device": E.device_name()}
    if path:
        with open(path, "w") as f:
            json.dump(out, f, indent=1)
    return out


def load(path: str = DEFAULT_PATH) -> dict | None:
    """The saved configuration (dict for Scene.set_launch_config) or None."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)["config"]


def main() -> int:
    """Tunes on the 64-scene batch (config D, depth 2, hipGraph replay) and saves
    the winner to eos_amd/tuned_launch.json; --single tunes one config C frame."""
    import argparse

    from . import scenes as S
    ap = argparse.ArgumentParser(description=main.__doc__.split("\n")[0])
    ap.add_argument("--single", action="store_true", help="tune on one config C frame instead of the batch")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=DEFAULT_PATH)
    args = ap.parse_args()
    if args.single:
        data = [S.config_c()]
        workload = "config C frame"
    else:
        data = S.config_d()
        workload = "config D: 64 soups, depth 2"
    sc = [E.Scene(d, leaf_max=4) for d in data]
    res = tune(sc, depth=2 if not args.single else data[0].depth, reps=args.reps, batched=not args.single,
               path=args.out, workload=workload)
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
