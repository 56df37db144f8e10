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


def _batch_scenes():
    """Scenes of every shape a batch can hold: soups of uneven sizes, index
    tie-breaks, a transformed instance scene, scenes of one, two and three
    triangles (leaf-only trees), and an empty scene."""
    out = list(scenes.config_d(n_scenes=10, scale=0.05))
    out += [CASES[k]() for k in ("duplicates", "tiny_5", "two", "config_a")]
    out.append(scenes.config_f(32, 24, instances=4, nu=16, nv=8))
    a = scenes.config_a(16, 16)
    out.append(scenes.SceneData(a.verts[5120:5121], a.normals[5120:5121], a.mat[5120:5121], a.materials, a.lights,
                                a.camera, 1, "one"))
    out.append(scenes.SceneData(a.verts[:0], a.normals[:0], a.mat[:0], a.materials, a.lights, a.camera, 1, "empty"))
    return out


@pytest.mark.parametrize("leaf_max,width", [(4, 4), (1, 4), (8, 4), (4, 2)])
def test_batched_build_equals_single_builds(leaf_max, width):
    """eos_build_lbvh_batch gives every scene the same Morton codes,
    permutation, binary and 4-wide nodes and triangle records as its own
    eos_build_lbvh_ex, bit for bit, and the same frame."""
    from eos_amd import _lib as E
    data = _batch_scenes()
    print(f"batched build of {len(data)} scenes ({sum(d.n_tri for d in data)} triangles), leaf_max={leaf_max} "
          f"width={width}", flush=True)
    single = [E.Scene(d, leaf_max=leaf_max, width=width) for d in data]
    batch = [E.Scene(d, build=False) for d in data]
    st = E.build_batch(batch, leaf_max=leaf_max, width=width)
    assert len({s["build_ms"] for s in st}) == 1
    for d, a, b in zip(data, single, batch):
        assert a.build_stats["num_nodes4"] == b.build_stats["num_nodes4"], d.name
        assert a.build_stats["root4"] == b.build_stats["root4"], d.name
        if d.n_tri == 0:
            continue
        for what in ("morton", "morton_sorted", "perm", "tris", "nodes", "nodes4"):
            if what == "nodes" and d.n_tri < 2:
                continue
            assert np.array_equal(_u32(a.debug_read(what)), _u32(b.debug_read(what))), (d.name, what)
        ra = a.render(want_hits=True)
        rb = b.render(want_hits=True)
        assert np.array_equal(ra["rgba"], rb["rgba"]) and np.array_equal(ra["hit"], rb["hit"]), d.name
    with pytest.raises(E.EosError) as ei:
        batch[0].debug_read("kchild")
    assert ei.value.status == E.EOS_ERR_STATE
    # rebuilt one by one afterwards, the scenes keep working as before
    batch[0].build(leaf_max, width)
    assert np.array_equal(_u32(batch[0].debug_read("kchild")), _u32(single[0].debug_read("kchild")))


def test_batched_build_full_config_d():
    """The 64 scenes of config D (7.45M triangles): one batched build equals
    64 single builds on the 4-wide nodes and triangle records; the batch renders
    identically."""
    from eos_amd import _lib as E
    data = scenes.config_d()
    single = [E.Scene(d, leaf_max=4) for d in data]
    batch = [E.Scene(d, build=False) for d in data]
    st = E.build_batch(batch, leaf_max=4)
    print(f"config D batched build: {st[0]['build_ms']:.2f} ms for {sum(d.n_tri for d in data)} triangles, "
          f"single builds {sum(s.build_stats['build_ms'] for s in single):.2f} ms", flush=True)
    for d, a, b in zip(data, single, batch):
        assert np.array_equal(_u32(a.debug_read("nodes4")), _u32(b.debug_read("nodes4"))), d.name
        assert np.array_equal(_u32(a.debug_read("tris")), _u32(b.debug_read("tris"))), d.name
    oa, _ = E.render_batch(single, depth=2)
    ob, _ = E.render_batch(batch, depth=2)
    for x, y in zip(oa, ob):
        assert np.array_equal(x, y)
