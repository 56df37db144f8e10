"""Autotuner for the traversal launch configuration.

Searches workgroup size, per-lane LDS stack depth, persistent waves per CU and
the refill batch (how many idle lanes a wave refills at once) on a given
workload, and persists the winner as JSON.  ``load`` returns a saved
configuration for ``Scene.set_launch_config``.
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
    "workgroup": (64, 128, 256),
    "stack_lds": (8, 16, 32),
    "waves_per_cu": (16, 32),
    "refill": (8, 32),
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


def tune(scenes, depth: int, reps: int = 3, grid=None, batch: bool = True, use_graph: bool = True,
         path: str | None = DEFAULT_PATH, workload: str = "") -> dict:
    """Times every configuration of ``grid`` and returns (and saves) the fastest.

    batch=True times eos_render_batch over ``scenes`` (the batch uses scenes[0]'s
    configuration); batch=False times eos_render of scenes[0]."""
    grid = grid or GRID
    keys = list(grid)
    results = []
    t0 = time.perf_counter()
    for vals in itertools.product(*(grid[k] for k in keys)):
        cfg = dict(zip(keys, vals))
        scenes[0].set_launch_config(**cfg)
        ms = _time_batch(scenes, depth, reps, use_graph) if batch else _time_single(scenes[0], depth, reps)
        results.append((ms, cfg))
    results.sort(key=lambda r: r[0])
    best_ms, best = results[0]
    scenes[0].set_launch_config(**best)
    out = {"config": best, "ms": best_ms, "worst_ms": results[-1][0], "tried": len(results),
           "seconds": time.perf_counter() - t0, "workload": workload, "batch": batch,
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
