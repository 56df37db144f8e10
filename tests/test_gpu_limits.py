"""Frames whose shadow rays outnumber 32-bit item ids.

Traversal launches number their items with 32-bit ints.  A chain frame's
deeper levels normally trace their shadow rays in one launch; when the
levels' surfaces times the light slots could pass 2^31 - 1 they go in several
launches of consecutive levels, and the frame must equal the same frame
rendered as row bands (each band small enough for one launch).  A frame
whose single level could pass the limit is rejected with EOS_ERR_LIMIT before
any launch."""
import numpy as np
import pytest

from eos_amd import scenes

pytestmark = pytest.mark.gpu


def _many_lights(width, height, n_lights, depth):
    a = scenes.config_a(width, height)
    rng = np.random.default_rng(5)
    pos = np.stack([rng.uniform(-4, 4, n_lights), rng.uniform(2, 6, n_lights), rng.uniform(-4, 4, n_lights)], 1)
    lights = scenes.make_lights(pos, colors=np.full((n_lights, 3), 1.0 / n_lights))
    mats = a.materials.copy()
    mats[1] = scenes.make_material((0.6, 0.6, 0.6), (0.2, 0.2, 0.2), 16, 0.3, True)   # mirror-ish plane
    return scenes.SceneData(a.verts, a.normals, a.mat, mats, lights, a.camera, depth, "many lights")


def test_deep_levels_split_into_launches_of_32bit_items():
    from eos_amd import _lib as E
    s = _many_lights(4096, 4096, 32, 4)      # a level: 16.8M pixels x 32 light slots = 2^29 items
    print("limits: 4096x4096, 32 lights, depth 4", flush=True)
    g = E.Scene(s, leaf_max=4)
    full = g.render(want_hits=True)
    st = full["stats"]
    print(f"  shadow rays {st['shadow_rays']}, launches {st['trace_shadow_launches']}", flush=True)
    assert st["trace_shadow_launches"] == 3     # level 0; levels 1-3 (3 x 2^29 < 2^31); level 4
    bands = [g.render_rows(r, r + 1024, want_hits=True) for r in range(0, 4096, 1024)]
    assert all(b["stats"]["trace_shadow_launches"] == 2 for b in bands)
    assert np.array_equal(np.concatenate([b["rgba"] for b in bands]), full["rgba"])
    assert np.array_equal(np.concatenate([b["hit"] for b in bands]), full["hit"])
    assert sum(b["stats"]["shadow_rays"] for b in bands) == st["shadow_rays"]
    assert sum(b["stats"]["secondary_rays"] for b in bands) == st["secondary_rays"]
    assert st["secondary_rays"] > 0


def test_a_level_beyond_32bit_items_is_rejected():
    from eos_amd import _lib as E
    s = _many_lights(8192, 8192, 32, 1)      # 67.1M pixels x 32 light slots = 2^31 items
    print("limits: 8192x8192, 32 lights (rejected)", flush=True)
    g = E.Scene(s, leaf_max=4)
    with pytest.raises(E.EosError) as ei:
        g.render()
    assert ei.value.status == E.EOS_ERR_LIMIT
    assert "2^31" in str(ei.value)
    # the scene stays usable
    small = g.render_rows(0, 16)
    assert small["rgba"].shape == (16, 8192, 4)
