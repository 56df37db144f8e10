"""Transparent materials on the GPU: ray trees (a hit may spawn a reflection
ray and the Fresnel-weighted second ray of src/RayTracer.cpp:361-408) vs the
oracle's recursion, which reproduces the reference's own glass images byte for
byte (tests/test_golden.py).

Bars as for the other frames: primary hit IDs on >= 99.99% of pixels, RGBA
within 1/255, float colour within 2e-5 where the primary hits agree, ray counts
exact (the reference traces the Fresnel ray twice with one result; both here
trace it once and count it once).
"""
import numpy as np
import pytest

from eos_amd import scenes
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

from tests.test_gpu_render import _compare  # noqa: E402

CASES = {
    "glass_96x72_d3": lambda: scenes.glass(96, 72, level=2),
    "glass_64x48_d5": lambda: scenes.glass(64, 48, level=2, depth=5),
    "glass_160x120_d3_l4": lambda: scenes.glass(160, 120, level=4),
    "glass_d0": lambda: scenes.glass(48, 32, level=2, depth=0),
}


@pytest.mark.parametrize("width", [4, 2])
@pytest.mark.parametrize("name", sorted(CASES))
def test_glass_parity(name, width):
    from eos_amd import _lib as E
    s = CASES[name]()
    print(f"glass case={name} n={s.n_tri} {s.width}x{s.height} depth={s.depth} width={width}", flush=True)
    g = E.Scene(s, leaf_max=4, width=width)
    go = g.render(want_rgb=True, want_hits=True)
    oo = po.Oracle(s).render()
    agree, d = _compare(s, go, oo, name)
    st, oc = go["stats"], oo["counts"]
    print(f"  agree={agree} maxdiff={d.max()} rays={st['primary_rays']}+{st['shadow_rays']}"
          f"+{st['secondary_rays']} hits={st['hits']}", flush=True)
    assert st["primary_rays"] == oc["primary"]
    assert st["shadow_rays"] == oc["shadow"]
    assert st["secondary_rays"] == oc["secondary"]
    assert st["hits"] == oc["hits"]
    if s.depth >= 2:
        assert st["secondary_rays"] > st["hits"] // 4      # trees, not chains: many second rays


def test_glass_aa_and_quad_walk():
    from eos_amd import _lib as E
    s = CASES["glass_96x72_d3"]()
    print("glass aa and quad walk", flush=True)
    g = E.Scene(s, leaf_max=4)
    go = g.render(want_rgb=True, aa_samples=2)
    oo = po.Oracle(s).render_aa(2)
    d = np.abs(go["rgba"].astype(int) - oo["rgba"].astype(int)).max(-1)
    assert (d <= 1).mean() >= 0.9999
    g.set_launch_config(coop=4)
    a = g.render(want_rgb=True, want_hits=True)
    g.set_launch_config(coop=0)
    b = g.render(want_rgb=True, want_hits=True)
    assert np.array_equal(a["hit"], b["hit"])
    assert np.array_equal(a["rgb"].view(np.uint32), b["rgb"].view(np.uint32))


def test_glass_batch_and_rules():
    from eos_amd import _lib as E
    print("glass batch and rules", flush=True)
    ss = [CASES["glass_96x72_d3"](), scenes.config_a(64, 48), CASES["glass_64x48_d5"]()]
    gs = [E.Scene(s, leaf_max=4) for s in ss]
    for use_graph in (False, True):
        outs, _ = E.render_batch(gs, depth=3, use_graph=use_graph)
        for k, o in enumerate(outs):
            assert np.array_equal(o, gs[k].render(depth=3)["rgba"]), (k, use_graph)
    g = gs[0]
    with pytest.raises(E.EosError) as ei:
        g.render(keep_records=True)
    assert ei.value.status == E.EOS_ERR_INVALID
    with pytest.raises(E.EosError) as ei:
        g.adjoint(np.zeros((72, 96, 3), np.float32), recompute=True)
    assert ei.value.status == E.EOS_ERR_INVALID
    # no transparent material left: chains again, same image as the oracle
    m = ss[0].materials.copy()
    m["transparent"] = 0
    g.set_materials(m)
    plain = scenes.SceneData(ss[0].verts, ss[0].normals, ss[0].mat, m, ss[0].lights, ss[0].camera, 3, "opaque")
    plain.ambient_index = ss[0].ambient_index
    assert np.array_equal(g.render()["rgba"], po.Oracle(plain).render()["rgba"]) or \
        np.abs(g.render()["rgba"].astype(int) - po.Oracle(plain).render()["rgba"]).max() <= 1


def _glass_instances(width, height, depth=3):
    """The glass scene's three spheres as one local-space icosphere placed by
    per-triangle transformations (rotation, non-uniform scale, one mirrored),
    over the world-space plane."""
    import math
    base = scenes.glass(width, height, level=2, depth=depth)
    sv, sn = scenes.icosphere(2)
    pv, pn = scenes.quad_grid(24, 4.0, -0.7)      # the glass scene's plane
    xforms = np.stack([scenes.affine((0.3, 1.0, 0.2), 0.7, (0.7, 0.6, 0.75), (-1.2, 0.0, 0.0)),
                       scenes.affine((1.0, 0.0, 0.4), 1.9, (0.7, 0.7, 0.55), (0.3, 0.0, -0.6)),
                       scenes.affine((0.0, 0.0, 1.0), math.pi, (-0.5, 0.5, 0.5), (1.4, -0.2, 0.4))])
    n = len(sv)
    verts = np.concatenate([sv, sv, sv, pv])
    normals = np.concatenate([sn, sn, sn, pn])
    mat = np.concatenate([np.full(n, 0, np.int32), np.full(n, 1, np.int32), np.full(n, 2, np.int32),
                          np.full(len(pv), 3, np.int32)])
    xi = np.concatenate([np.repeat(np.arange(3, dtype=np.int32), n), np.full(len(pv), -1, np.int32)])
    s = scenes.SceneData(verts, normals, mat, base.materials, base.lights, base.camera, depth,
                         "glass instances", {}, xforms, xi)
    s.ambient_index = base.ambient_index
    return s


@pytest.mark.parametrize("depth", [3, 5])
def test_glass_with_transforms(depth):
    """Ray trees over transformed triangles: the hit point through M, the normal
    through transpose(inverse), both Fresnel children from the world-space hit."""
    from eos_amd import _lib as E
    s = _glass_instances(96, 72, depth)
    print(f"glass with transforms depth={depth} n={s.n_tri}", flush=True)
    g = E.Scene(s, leaf_max=4)
    go = g.render(want_rgb=True, want_hits=True)
    oo = po.Oracle(s).render()
    agree, d = _compare(s, go, oo, "glass_xf")
    st, oc = go["stats"], oo["counts"]
    print(f"  agree={agree} maxdiff={d.max()} rays={st['primary_rays']}+{st['shadow_rays']}+{st['secondary_rays']}",
          flush=True)
    assert st["primary_rays"] == oc["primary"]
    assert st["shadow_rays"] == oc["shadow"]
    assert st["secondary_rays"] == oc["secondary"]
    assert st["secondary_rays"] > 0
    inst = s.xform_index[go["hit"][go["hit"] >= 0]]
    assert set(np.unique(inst[inst >= 0])) == {0, 1, 2}
