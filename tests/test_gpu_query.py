"""Ray queries on the GPU (eos_trace_rays) against the reference itself.

The probe rays of tests/golden/reference_vectors.npz were answered by the
reference's own Scene::findMinDistanceIntersectionBVH,
Scene::shadowRayVisibilityBVH and Scene::findMinDistanceIntersectionLinear
(compiled from its sources by oracle/ref.mk, driven by oracle/ref_harness.cpp).
The HIP traversal must give the same hit flag, point, normal and material bit
for bit.  The BVH queries leave out the rays tests/test_golden.py leaves out
(a zero direction component makes the reference's slab test NaN; a hit closer
than the reference's 1e-3 box threshold); the linear query has no box test, so
every ray counts.
"""
import numpy as np
import pytest

from tests.golden.make_golden import golden_scenes, scene_hash
from tests.test_golden import GOLD, bunny_scene
from eos_amd import scenes
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD, allow_pickle=False)


@pytest.mark.parametrize("width", [4, 2])
@pytest.mark.parametrize("name", sorted(golden_scenes()))
def test_queries_match_reference(gold, name, width):
    from eos_amd import _lib as E
    s = golden_scenes()[name]()
    assert bytes(gold[f"{name}/hash"]).decode() == scene_hash(s)
    print(f"queries scene={name} n={s.n_tri} width={width}", flush=True)
    g = E.Scene(s, leaf_max=4, width=width)
    o, d = gold[f"{name}/ray_o"], gold[f"{name}/ray_d"]
    r = g.trace_rays(o, d, "nearest")
    hit = r["triangle"] >= 0
    assert np.array_equal(hit, r["hit"] == 1)
    near = hit & (r["t"] < 1e-3)
    assert near.sum() <= 2
    keep = ~(d == 0).any(1) & ~near
    assert np.array_equal(hit[keep], gold[f"{name}/hit"][keep].astype(bool))
    h = hit & keep
    assert h.sum() > 20
    assert np.array_equal(r["point"][h], gold[f"{name}/point"][h])
    assert np.array_equal(r["normal"][h], gold[f"{name}/normal"][h])
    assert np.array_equal(r["material"][h], gold[f"{name}/mat"][h])
    assert np.array_equal(r["material"][hit], s.mat[r["triangle"][hit]])
    assert (r["material"][~hit] == -1).all() and (r["point"][~hit] == 0).all()
    occ = g.trace_rays(o, d, "occluded")
    assert np.array_equal(occ[keep], gold[f"{name}/occluded"][keep].astype(bool))
    lin = g.trace_rays(o, d, "linear")
    assert np.array_equal(lin["hit"] == 1, gold[f"{name}/lin_hit"].astype(bool))
    f = lin["hit"] == 1
    assert np.array_equal(lin["point"][f], gold[f"{name}/lin_point"][f])
    assert np.array_equal(lin["normal"][f], gold[f"{name}/lin_normal"][f])
    k = lin["triangle"] >= 0
    assert np.array_equal(s.materials["kd"][lin["material"][k]], gold[f"{name}/lin_kd"][k])
    # the two nearest-hit queries find the same triangle away from the exclusions
    assert np.array_equal(lin["triangle"][keep], r["triangle"][keep])


def test_queries_on_the_reference_driver_scene(gold):
    """The reference driver's bunny (src/main.cpp:330-411): the linear query
    reproduces the reference on all 14,400 sampled primary rays and the shadow
    rays of their hits; the BVH queries find the same hits and points."""
    from eos_amd import _lib as E
    s = bunny_scene()
    assert bytes(gold["bunny/hash"]).decode() == scene_hash(s)
    print(f"queries on the driver scene: {s.n_tri} triangles", flush=True)
    g = E.Scene(s, leaf_max=4)
    o, d = gold["bunny/ray_o"], gold["bunny/ray_d"]
    lin = g.trace_rays(o, d, "linear")
    f = lin["hit"] == 1
    assert np.array_equal(f, gold["bunny/lin_hit"].astype(bool))
    assert np.array_equal(lin["point"][f], gold["bunny/lin_point"][f])
    assert np.array_equal(lin["normal"][f], gold["bunny/lin_normal"][f])
    assert (lin["material"][f] == 0).all()
    so, sd = gold["bunny/shadow_o"], gold["bunny/shadow_d"]
    assert np.array_equal(g.trace_rays(so, sd, "linear")["hit"] == 1, gold["bunny/shadow_hit"].astype(bool))
    assert np.array_equal(g.trace_rays(so, sd, "occluded"), gold["bunny/shadow_hit"].astype(bool))
    r = g.trace_rays(o, d, "nearest")
    assert np.array_equal(r["triangle"] >= 0, f)
    assert np.array_equal(r["triangle"], lin["triangle"])
    assert np.array_equal(r["point"][f], lin["point"][f])
    assert np.abs(r["normal"][f] - lin["normal"][f]).max() <= 2.0 ** -22


def test_random_rays_match_oracle_with_rotations():
    """Rotated, scaled and mirrored instances (no reference golden covers
    rotations, see tests/test_golden.py): GPU queries equal the oracle's on random
    rays in and around the scene, in host and in device memory."""
    import torch
    from eos_amd import _lib as E
    s = scenes.config_f(64, 48, instances=9, nu=24, nv=12)
    rng = np.random.default_rng(11)
    v = s.verts.reshape(-1, 3)
    lo, hi = v.min(0) - 1.0, v.max(0) + 1.0
    n = 3000
    o = (lo + (hi - lo) * rng.random((n, 3))).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d[: n // 3] *= 3.5            # directions of any length: t is in units of the direction
    print(f"random rays n={n} tris={s.n_tri}", flush=True)
    g = E.Scene(s, leaf_max=4)
    orc = po.Oracle(s)
    r = g.trace_rays(o, d, "nearest")
    b = orc.nearest(o, d)
    assert np.array_equal(r["triangle"], b["tri"])
    h = b["tri"] >= 0
    assert h.sum() > 300
    assert np.array_equal(r["t"][h], b["t"][h])
    assert np.array_equal(r["point"][h], b["point"][h]) and np.array_equal(r["normal"][h], b["normal"][h])
    assert np.array_equal(g.trace_rays(o, d, "occluded"), orc.nearest(o, d, any_hit=True))
    lin = g.trace_rays(o, d, "linear")
    ol = orc.nearest_linear(o, d)
    assert np.array_equal(lin["hit"] == 1, ol["found"]) and np.array_equal(lin["triangle"], ol["tri"])
    k = ol["tri"] >= 0
    assert np.array_equal(lin["point"][k], ol["point"][k]) and np.array_equal(lin["normal"][k], ol["normal"][k])
    # device memory: the same records
    to, td = torch.from_numpy(o).cuda(), torch.from_numpy(d).cuda()
    hits = torch.zeros(n * 12, dtype=torch.int32, device="cuda")
    occ = torch.zeros(n, dtype=torch.uint8, device="cuda")
    g.trace_rays_device(n, to.data_ptr(), td.data_ptr(), "nearest", hits_ptr=hits.data_ptr())
    g.trace_rays_device(n, to.data_ptr(), td.data_ptr(), "occluded", occluded_ptr=occ.data_ptr())
    torch.cuda.synchronize()
    got = hits.cpu().numpy().view(E.RAY_HIT)
    assert np.array_equal(got.view(np.uint8), r.view(np.uint8))
    assert np.array_equal(occ.cpu().numpy().astype(bool), orc.nearest(o, d, any_hit=True))


def test_many_rays_and_edge_cases():
    """A launch far larger than the persistent grid (per-lane walk), one ray
    (quad walk), zero rays, and rays that start on a surface or miss everything."""
    from eos_amd import _lib as E
    s = scenes.config_c(96, 64, nu=256, nv=128)
    g = E.Scene(s, leaf_max=4)
    orc = po.Oracle(s)
    rng = np.random.default_rng(3)
    n = 600_000
    o = (rng.random((n, 3)) * 4.0 - 2.0).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    print(f"many rays n={n}", flush=True)
    r = g.trace_rays(o, d, "nearest")
    sub = rng.choice(n, 20000, replace=False)
    b = orc.nearest(o[sub], d[sub])
    assert np.array_equal(r["triangle"][sub], b["tri"])
    occ = g.trace_rays(o, d, "occluded")
    assert np.array_equal(occ[sub], orc.nearest(o[sub], d[sub], any_hit=True))
    assert np.array_equal(occ, r["triangle"] >= 0)    # occluded exactly when some triangle is hit at t >= 0
    one = g.trace_rays(o[:1], d[:1], "nearest")
    assert one.view(np.uint8).tobytes() == r[:1].view(np.uint8).tobytes()
    assert len(g.trace_rays(o[:0], d[:0], "nearest")) == 0
    away = g.trace_rays(np.full((4, 3), 50.0, np.float32), np.ones((4, 3), np.float32), "nearest")
    assert (away["hit"] == 0).all() and (away["triangle"] == -1).all() and np.isinf(away["t"]).all()
