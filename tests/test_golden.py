"""The oracle against vectors produced by the reference's own CPU renderer.

tests/golden/reference_vectors.npz was written by tests/golden/make_golden.py
from the reference's RayTracer / Scene compiled in place (oracle/ref.mk).
Expected agreement is exact: same float operations, same libm.  The one
documented exception is a reference defect: when a ray direction component is
exactly 0 the reference's slab test computes (0 - 0) * inf = NaN for boxes
with a bound at that coordinate and culls them (src/Box.cu:228-256 with the
1/d inverse of src/Ray.h:47), so such rays miss; those pixels are excluded.
"""
import os

import numpy as np
import pytest

from tests.golden.make_golden import golden_scenes, scene_hash
from oracle import pyoracle as po

GOLD = os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD, allow_pickle=False)


def _zero_component_pixels(s):
    """Pixels whose primary direction has an exactly-zero component."""
    cam = np.array(s.camera, copy=True)
    W, H = s.width, s.height
    lib = po.lib()
    mask = np.zeros((H, W), bool)
    d = np.zeros(3, np.float32)
    for y in range(H):
        for x in range(W):
            lib.orc_camera_ray(cam.ctypes.data, x, y, d.ctypes.data)
            mask[y, x] = bool((d == 0).any())
    return mask


@pytest.mark.parametrize("name", sorted(golden_scenes()))
def test_render_matches_reference(gold, name):
    s = golden_scenes()[name]()
    assert bytes(gold[f"{name}/hash"]).decode() == scene_hash(s), "scene generator changed"
    ref = gold[f"{name}/rgba"]
    out = po.Oracle(s).render()["rgba"]
    excl = _zero_component_pixels(s)
    diff = np.abs(ref.astype(int) - out.astype(int)).max(-1)
    assert (diff[~excl] == 0).all(), f"{(diff[~excl] > 0).sum()} pixels differ"
    assert ref.shape == out.shape


@pytest.mark.parametrize("name", sorted(golden_scenes()))
def test_nearest_and_occlusion_match_reference(gold, name):
    s = golden_scenes()[name]()
    o, d = gold[f"{name}/ray_o"], gold[f"{name}/ray_d"]
    keep = ~(d == 0).any(1)
    orc = po.Oracle(s)
    r = orc.nearest(o, d)
    hit = r["tri"] >= 0
    assert np.array_equal(hit[keep], gold[f"{name}/hit"][keep].astype(bool))
    h = hit & keep
    assert np.array_equal(r["point"][h], gold[f"{name}/point"][h])
    assert np.array_equal(r["normal"][h], gold[f"{name}/normal"][h])
    assert np.array_equal(s.mat[r["tri"][h]], gold[f"{name}/mat"][h])
    occ = orc.nearest(o, d, any_hit=True)
    assert np.array_equal(occ[keep], gold[f"{name}/occluded"][keep].astype(bool))
