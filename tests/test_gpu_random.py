"""Seeded random scenes through every feature at once, against the oracle.

Each seed draws a small scene: a soup of triangles of mixed sizes and an
instanced icosphere (per-triangle transformations with rotation, non-uniform
scale and sometimes a mirror), random materials (reflective, transparent with
a refractive index, both, neither), one to four lights, a random camera and
image size, bounce depth 0 to 4, tree width and leaf size, and anti-aliasing
on some seeds.  Bars as for the fixed scenes: primary hits on >= 99.99% of
pixels, RGBA within 1/255 where they agree, ray counts exact."""
import math

import numpy as np
import pytest

from eos_amd import scenes
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

from tests.test_gpu_render import _compare  # noqa: E402

SEEDS = list(range(64))


def _random_scene(seed):
    rng = np.random.default_rng(1000 + seed)
    n_soup = int(rng.integers(20, 1500))
    cen = rng.uniform(-1.0, 1.0, (n_soup, 1, 3))
    size = np.exp(rng.uniform(math.log(0.02), math.log(0.4), (n_soup, 1, 1)))
    tri = cen + size * rng.normal(size=(n_soup, 3, 3))
    fn = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    fn /= np.maximum(np.linalg.norm(fn, axis=1, keepdims=True), 1e-12)
    soup_n = np.repeat(fn[:, None, :], 3, axis=1)
    sv, sn = scenes.icosphere(int(rng.integers(1, 3)))
    n_inst = int(rng.integers(1, 4))
    xforms = []
    for k in range(n_inst):
        scale = rng.uniform(0.2, 0.6, 3)
        if rng.random() < 0.3:
            scale[int(rng.integers(0, 3))] *= -1.0           # a mirrored instance
        xforms.append(scenes.affine(rng.normal(size=3), rng.uniform(0, 2 * math.pi), scale,
                                    rng.uniform(-1.2, 1.2, 3)))
    verts = np.concatenate([tri] + [sv] * n_inst).astype(np.float32)
    normals = np.concatenate([soup_n] + [sn] * n_inst).astype(np.float32)
    xi = np.concatenate([np.full(n_soup, -1, np.int32)] +
                        [np.full(len(sv), k, np.int32) for k in range(n_inst)])
    n_mat = int(rng.integers(1, 5))
    mats = []
    for k in range(n_mat):
        refl, transp = bool(rng.random() < 0.5), bool(rng.random() < 0.35)
        mats.append(scenes.make_material(rng.uniform(0.05, 0.9, 3), rng.uniform(0.0, 0.8, 3),
                                         int(rng.integers(1, 96)), float(rng.uniform(0.1, 0.8)) if refl else 0.0,
                                         refl, transp, float(rng.uniform(1.05, 1.8))))
    mats = np.stack(mats)
    mat = rng.integers(0, n_mat, len(verts)).astype(np.int32)
    n_light = int(rng.integers(1, 5))
    lights = scenes.make_lights(rng.uniform(-3, 3, (n_light, 3)) + np.array([0.0, 3.0, 0.0]),
                                rng.uniform(0.2, 1.0, (n_light, 3)))
    W, H = int(rng.integers(8, 72)), int(rng.integers(8, 56))
    eye = rng.normal(size=3)
    eye = 3.2 * eye / np.linalg.norm(eye)
    cam = scenes.make_camera(eye, rng.uniform(-0.3, 0.3, 3), float(rng.uniform(35, 75)), W, H)
    depth = int(rng.integers(0, 5))
    use_xf = rng.random() < 0.75
    s = scenes.SceneData(verts, normals, mat, mats, lights, cam, depth, f"random {seed}", {},
                         np.stack(xforms) if use_xf else None, xi if use_xf else None)
    s.ambient_index = float(rng.choice([1.0, 1.0002, 1.1]))
    opts = {"leaf_max": int(rng.integers(1, 9)), "width": int(rng.choice([2, 4])),
            "aa": int(rng.choice([0, 0, 0, 2, 3]))}
    return s, opts


@pytest.mark.parametrize("seed", SEEDS)
def test_random_scene_matches_oracle(seed):
    from eos_amd import _lib as E
    s, o = _random_scene(seed)
    print(f"random seed={seed} n={s.n_tri} {s.width}x{s.height} depth={s.depth} lights={len(s.lights)} "
          f"xf={s.xforms is not None} transparent={int(s.materials['transparent'].sum())} {o}", flush=True)
    g = E.Scene(s, leaf_max=o["leaf_max"], width=o["width"])
    if o["aa"]:
        go = g.render(want_rgb=True, aa_samples=o["aa"])
        oo = po.Oracle(s).render_aa(o["aa"])
        d = np.abs(go["rgba"].astype(int) - oo["rgba"].astype(int)).max(-1)
        assert (d <= 1).mean() >= 0.999, f"seed {seed}: {(d > 1).sum()} pixels off by more than 1"
        oc = oo["counts"]
    else:
        go = g.render(want_rgb=True, want_hits=True)
        oo = po.Oracle(s).render()
        _compare(s, go, oo, f"random{seed}")
        oc = oo["counts"]
    st = go["stats"]
    assert st["primary_rays"] == oc["primary"]
    assert st["shadow_rays"] == oc["shadow"]
    assert st["secondary_rays"] == oc["secondary"]


@pytest.mark.parametrize("group", range(4))
def test_random_batch_equals_single_frames(group):
    """Eight random scenes (built at one tree width) in one batched frame, with
    and without graph replay: every image equals that scene's own frame."""
    from eos_amd import _lib as E
    ss = [_random_scene(100 + 8 * group + k)[0] for k in range(8)]
    depth = 1 + group % 3
    print(f"random batch group={group} depth={depth} sizes={[s.n_tri for s in ss]}", flush=True)
    gs = [E.Scene(s, leaf_max=4, width=4) for s in ss]
    singles = [g.render(depth=depth)["rgba"] for g in gs]
    for use_graph in (False, True):
        outs, st = E.render_batch(gs, depth=depth, use_graph=use_graph)
        for k, (o, want) in enumerate(zip(outs, singles)):
            assert np.array_equal(o, want), (group, k, use_graph)


@pytest.mark.parametrize("seed", range(8))
def test_random_bands_stitch(seed):
    """A random scene rendered in random row bands equals its whole frame."""
    from eos_amd import _lib as E
    s, o = _random_scene(200 + seed)
    rng = np.random.default_rng(seed)
    H = s.height
    cuts = np.unique(np.concatenate([[0, H], rng.integers(1, H, int(rng.integers(1, 5)))]))
    print(f"random bands seed={seed} {s.width}x{H} cuts={cuts.tolist()} aa={o['aa']}", flush=True)
    g = E.Scene(s, leaf_max=o["leaf_max"], width=o["width"])
    full = g.render(aa_samples=o["aa"])["rgba"]
    bands = np.concatenate([g.render_rows(int(a), int(b), aa_samples=o["aa"])["rgba"] for a, b in zip(cuts, cuts[1:])])
    assert np.array_equal(bands, full)


@pytest.mark.parametrize("seed", range(8))
def test_random_adjoint_matches_fp64_forward_mode(seed):
    """Gradients of the L2 image loss on random opaque scenes (transformations
    on most seeds) against the oracle's FP64 forward mode over the same hits."""
    from eos_amd import _lib as E
    s, o = _random_scene(300 + seed)
    s.materials["transparent"] = 0
    if s.depth == 0:
        s.depth = 1
    print(f"random adjoint seed={seed} n={s.n_tri} {s.width}x{s.height} depth={s.depth} "
          f"xf={s.xforms is not None}", flush=True)
    g = E.Scene(s, leaf_max=o["leaf_max"], width=o["width"])
    tm, tl = scenes.perturb_params(s, 77 + seed)
    g.set_materials(tm)
    g.set_lights(tl)
    target = g.render(want_rgb=True)["rgb"]
    g.set_materials(s.materials)
    g.set_lights(s.lights)
    fwd = g.render(want_rgb=True, want_hits=True, keep_records=True)
    gm, gl, _ = g.adjoint(2.0 * (fwd["rgb"] - target))
    out = po.Oracle(s).render(want_recs=True)
    assert np.array_equal(out["hit"], fwd["hit"])
    prm = po.params_vector(s.materials, s.lights)
    ref = po.grad_fwd64(out["recs"], s.camera["position"], len(s.materials), len(s.lights), prm, target)
    got = np.concatenate([gm.ravel(), gl.ravel()])
    scale = np.abs(ref).max()
    if scale == 0.0:
        assert np.abs(got).max() == 0.0
        return
    err = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-3 * scale)
    print(f"  max rel err {err.max():.2e}", flush=True)
    assert err.max() <= 1e-3


@pytest.mark.parametrize("seed", range(6))
def test_random_frames_in_flight(seed):
    """Five cameras around a random scene rendered as frames in flight equal
    five single renders, each with its own camera."""
    import torch
    from eos_amd import _lib as E
    s, o = _random_scene(400 + seed)
    rng = np.random.default_rng(seed)
    cams = []
    for k in range(5):
        eye = rng.normal(size=3)
        cams.append(scenes.make_camera(3.0 * eye / np.linalg.norm(eye), (0.0, 0.0, 0.0), 55.0, s.width, s.height))
    cams = np.stack(cams)
    print(f"random frames seed={seed} n={s.n_tri} {s.width}x{s.height} depth={s.depth}", flush=True)
    g = E.Scene(s, leaf_max=o["leaf_max"], width=o["width"])
    outs = [torch.empty(s.width * s.height * 4, dtype=torch.uint8, device="cuda:0") for _ in cams]
    g.render_frames(cams, [t.data_ptr() for t in outs])
    for k, c in enumerate(cams):
        want = g.render(camera=c)["rgba"]
        assert np.array_equal(outs[k].cpu().numpy().reshape(s.height, s.width, 4), want), k
