"""Batched rendering: many scenes in one launch per pipeline stage.

Every ray of a batched frame runs the same arithmetic as in a single-scene
frame, so each scene's image must equal its own eos_render output byte for
byte, and ray counts must add up.  The hipGraph path is replayed several
times with other GPU work in between; every replay is checked.
"""
import numpy as np
import pytest

from eos_amd import scenes
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch():
    from eos_amd import _lib as E
    data = scenes.config_d(n_scenes=12, scale=0.02)
    sc = [E.Scene(d, leaf_max=4) for d in data]
    return data, sc


def test_batch_equals_single_renders(batch):
    from eos_amd import _lib as E
    data, sc = batch
    print("batch of", len(sc), "scenes:", [d.n_tri for d in data], flush=True)
    singles = [s.render(depth=2) for s in sc]
    outs, st = E.render_batch(sc, depth=2)
    for d, s1, o in zip(data, singles, outs):
        assert np.array_equal(o, s1["rgba"]), d.name
    rays = sum(r["stats"]["primary_rays"] + r["stats"]["shadow_rays"] + r["stats"]["secondary_rays"] for r in singles)
    assert st["rays"] == rays
    # and against the oracle for a couple of scenes
    for k in (0, 5):
        ref = po.Oracle(data[k]).render(depth=2)
        assert (np.abs(outs[k].astype(int) - ref["rgba"].astype(int)) <= 1).mean() >= 0.9999


def test_batch_graph_replays(batch):
    from eos_amd import _lib as E
    data, sc = batch
    ref, _ = E.render_batch(sc, depth=2, use_graph=False)
    for rep in range(4):
        outs, st = E.render_batch(sc, depth=2, use_graph=True)
        print(f"graph replay {rep}: {st['render_ms']:.3f} ms", flush=True)
        for a, b in zip(outs, ref):
            assert np.array_equal(a, b)
        sc[rep % len(sc)].render(depth=2)          # eager work between replays


def test_batch_with_launch_configs(batch):
    from eos_amd import _lib as E
    data, sc = batch
    ref, _ = E.render_batch(sc, depth=2)
    for wg, stack, batch_items, packets, coop in ((64, 8, 64, 1, 4), (128, 32, 1024, 1, 0), (256, 16, 16, 0, 1)):
        sc[0].set_launch_config(wg, stack, 16, 8, batch_items, packets, coop)
        outs, _ = E.render_batch(sc, depth=2, use_graph=True)
        for a, b in zip(outs, ref):
            assert np.array_equal(a, b)
    sc[0].set_launch_config()


def test_batch_follows_material_changes():
    """A material turned transparent (or back) between two batched frames: the
    batch is reshaped (ray trees instead of chains) and every image still equals
    the scene's single render, with and without graph replay."""
    from eos_amd import _lib as E
    data = scenes.config_d(n_scenes=4, scale=0.02)
    sc = [E.Scene(d, leaf_max=4) for d in data]
    for use_graph in (False, True):
        base, _ = E.render_batch(sc, depth=2, use_graph=use_graph)
        for d, s, o in zip(data, sc, base):
            assert np.array_equal(o, s.render(depth=2)["rgba"]), d.name
        m = data[1].materials.copy()
        m["transparent"][0] = 1
        m["refractive_index"][0] = 1.4
        sc[1].set_materials(m)
        outs, _ = E.render_batch(sc, depth=2, use_graph=use_graph)
        glass = sc[1].render(depth=2)["rgba"]
        assert not np.array_equal(glass, base[1])
        print(f"use_graph={use_graph}: scene 1 changed on {int((glass != base[1]).any(-1).sum())} pixels", flush=True)
        for d, s, o in zip(data, sc, outs):
            assert np.array_equal(o, s.render(depth=2)["rgba"]), d.name
        sc[1].set_materials(data[1].materials)
        outs, _ = E.render_batch(sc, depth=2, use_graph=use_graph)
        for o, b in zip(outs, base):
            assert np.array_equal(o, b)
