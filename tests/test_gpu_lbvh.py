"""GPU LBVH build vs the CPU Karras restatement: bit-exact at every stage."""
import numpy as np
import pytest

from eos_amd import scenes
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _dup_scene():
    # many coincident centroids -> equal Morton codes -> index tie-breaks everywhere
    v, n = scenes.quad_grid(16, 1.0, 0.0)
    v = np.concatenate([v, v, v[:100], v[:7]])
    n = np.concatenate([n, n, n[:100], n[:7]])
    s = scenes.config_a(32, 32)
    return scenes.SceneData(v, n, np.zeros(len(v), np.int32), s.materials[:1], s.lights, s.camera, 1, "dup")


def _sphere():
    s = scenes.config_c(96, 54, nu=96, nv=48)
    return s


CASES = {
    "config_a": lambda: scenes.config_a(64, 64),
    "soup_30k": lambda: scenes.config_b(64, 48, n=30000),
    "duplicates": _dup_scene,
    "displaced_sphere": _sphere,
    "tiny_5": lambda: scenes.config_b(16, 16, n=5),
    "two": lambda: scenes.config_b(16, 16, n=2),
}


def _u32(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.parametrize("leaf_max", [1, 4, 8])
@pytest.mark.parametrize("name", sorted(CASES))
def test_lbvh_bit_exact(name, leaf_max):
    from eos_amd import _lib as E
    s = CASES[name]()
    print(f"lbvh case={name} n={s.n_tri} leaf_max={leaf_max}", flush=True)
    g = E.Scene(s, leaf_max=leaf_max)
    ref = po.lbvh_build(s.verts, leaf_max)
    codes, _ = po.morton_codes(s.verts)
    assert np.array_equal(g.debug_read("morton"), codes)
    assert np.array_equal(g.debug_read("morton_sorted"), ref["codes"])
    assert np.array_equal(g.debug_read("perm"), ref["perm"])
    assert g.build_stats["root_ref"] == ref["root"]
    assert np.array_equal(_u32(g.debug_read("tris"))[:, :11], _u32(ref["tris"])[:, :11])
    assert g.build_stats["num_nodes4"] == len(ref["nodes4"])
    assert g.build_stats["root4"] == ref["root4"]
    if len(ref["nodes4"]):
        assert np.array_equal(_u32(g.debug_read("nodes4"))[:, :28], _u32(ref["nodes4"])[:, :28])
    if s.n_tri > 1:
        assert np.array_equal(g.debug_read("kchild"), ref["kchild"])
        assert np.array_equal(g.debug_read("kparent"), ref["kparent"])
        assert np.array_equal(_u32(g.debug_read("ibox")), _u32(ref["ibox"]))
        assert np.array_equal(_u32(g.debug_read("nodes"))[:, :14], _u32(ref["nodes"])[:, :14])


def test_lbvh_large_sorted_and_consistent():
    """1M soup: full-size structural invariants (sortedness, permutation, every
    internal node reached once, children ranges partition the parent's)."""
    from eos_amd import _lib as E
    s = scenes.config_b(32, 32, n=1_000_000)
    print("lbvh 1M soup", flush=True)
    g = E.Scene(s, leaf_max=1)
    codes = g.debug_read("morton_sorted")
    perm = g.debug_read("perm")
    assert np.all(codes[1:] >= codes[:-1])
    assert np.array_equal(np.sort(perm), np.arange(s.n_tri))
    ref_codes, _ = po.morton_codes(s.verts)
    assert np.array_equal(ref_codes[perm], codes)
    # stability: equal codes keep ascending original index
    eq = codes[1:] == codes[:-1]
    assert np.all(perm[1:][eq] > perm[:-1][eq])
    kc = g.debug_read("kchild")
    kp = g.debug_read("kparent")
    n = s.n_tri
    internal = kc[kc >= 0]
    leaves = ~kc[kc < 0]
    assert len(internal) == n - 2 and len(np.unique(internal)) == n - 2
    assert len(leaves) == n and np.array_equal(np.sort(leaves), np.arange(n))
    assert kp[0] == -1
    idx = np.arange(n - 1)
    for c in (0, 1):
        ch = kc[:, c]
        par_of = np.where(ch >= 0, kp[np.maximum(ch, 0)], kp[n - 1 + np.maximum(~ch, 0)])
        assert np.array_equal(par_of, idx)
