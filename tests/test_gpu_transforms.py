"""Per-triangle transformations on the GPU (eos_set_transforms) vs the oracle.

The oracle's transformation path is pinned against the reference's own CPU
renderer (axis-aligned matrices, tests/test_golden.py), against glm::inverse
bit for bit, and against brute force and pre-transformed geometry for general
rotations (tests/test_oracle.py).  Here the GPU build and frames are held to
the same bars as the world-space scenes: the LBVH bit-exact, hit IDs on
>= 99.99% of pixels, RGBA within 1/255, ray counts and walk counters exact.
"""
import numpy as np
import pytest

from eos_amd import scenes
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

from tests.test_gpu_render import _compare  # noqa: E402


def _u32(a):
    return np.ascontiguousarray(a).view(np.uint32)


CASES = {
    "inst8_rot": lambda: scenes.config_f(96, 72, instances=8, nu=24, nv=12),
    "inst8_axis": lambda: scenes.config_f(96, 72, instances=8, nu=24, nv=12, rotate=False),
    "inst64_rot": lambda: scenes.config_f(128, 96, instances=64, nu=16, nv=8),
}


@pytest.mark.parametrize("leaf_max", [1, 4])
@pytest.mark.parametrize("name", sorted(CASES))
def test_lbvh_bit_exact_with_transforms(name, leaf_max):
    from eos_amd import _lib as E
    s = CASES[name]()
    print(f"transform lbvh case={name} n={s.n_tri} leaf_max={leaf_max}", flush=True)
    g = E.Scene(s, leaf_max=leaf_max)
    ref = po.lbvh_build(s.verts, leaf_max, s.xforms, s.xform_index)
    assert np.array_equal(g.debug_read("morton_sorted"), ref["codes"])
    assert np.array_equal(g.debug_read("perm"), ref["perm"])
    assert np.array_equal(_u32(g.debug_read("ibox")), _u32(ref["ibox"]))
    assert np.array_equal(_u32(g.debug_read("tris"))[:, :11], _u32(ref["tris"])[:, :11])
    assert np.array_equal(_u32(g.debug_read("nodes4"))[:, :28], _u32(ref["nodes4"])[:, :28])
    # device transformation rows: inverse, M, transpose(inverse), glm's inverse bit for bit
    rows = g.debug_read("xforms").reshape(-1, 9, 4)
    for k, m in enumerate(s.xforms):
        inv = po.mat4_inverse(m).reshape(4, 4)          # [column][row]
        M = m.reshape(4, 4)
        assert np.array_equal(_u32(rows[k, 0:3]), _u32(inv[:, :3].T))
        assert np.array_equal(_u32(rows[k, 3:6]), _u32(M[:, :3].T))
        assert np.array_equal(_u32(rows[k, 6:9]), _u32(inv[:3, :]))


@pytest.mark.parametrize("width", [4, 2])
@pytest.mark.parametrize("name", sorted(CASES))
def test_render_parity_with_transforms(name, width):
    from eos_amd import _lib as E
    s = CASES[name]()
    print(f"transform render case={name} n={s.n_tri} {s.width}x{s.height} width={width}", flush=True)
    g = E.Scene(s, leaf_max=4, width=width)
    go = g.render(want_rgb=True, want_hits=True)
    oo = po.Oracle(s).render()
    agree, d = _compare(s, go, oo, name)
    st, oc = go["stats"], oo["counts"]
    print(f"  agree={agree} maxdiff={d.max()} rays={st['primary_rays']}+{st['shadow_rays']}"
          f"+{st['secondary_rays']}", flush=True)
    assert st["primary_rays"] == oc["primary"]
    assert st["shadow_rays"] == oc["shadow"]
    assert st["secondary_rays"] == oc["secondary"]
    assert st["hits"] == oc["hits"]
    inst = s.xform_index[go["hit"][go["hit"] >= 0]]
    assert len(np.unique(inst[inst >= 0])) >= 6


@pytest.mark.parametrize("coop", [0, 4])
def test_walk_counters_with_transforms(coop):
    """Node visits and triangle tests equal the restatement of the same walk,
    per-lane and quad-cooperative."""
    from eos_amd import _lib as E
    s = CASES["inst8_rot"]()
    print(f"transform counters coop={coop}", flush=True)
    g = E.Scene(s, leaf_max=4)
    g.set_launch_config(coop=coop)
    go = g.render(count_visits=True, want_hits=True)
    b = po.lbvh_build(s.verts, 4, s.xforms, s.xform_index)
    oo = po.render_lbvh(s, b, width=4)
    assert np.array_equal(go["hit"], oo["hit"])
    assert go["stats"]["node_visits"] == oo["counts"]["node_visits"]
    assert go["stats"]["tri_tests"] == oo["counts"]["tri_tests"]


def test_quad_and_per_lane_walks_agree_with_transforms():
    from eos_amd import _lib as E
    s = CASES["inst64_rot"]()
    print("transform quad vs per-lane", flush=True)
    g = E.Scene(s, leaf_max=4)
    g.set_launch_config(coop=4)
    a = g.render(want_rgb=True, want_hits=True)
    g.set_launch_config(coop=0)
    b = g.render(want_rgb=True, want_hits=True)
    assert np.array_equal(a["hit"], b["hit"])
    assert np.array_equal(a["rgb"].view(np.uint32), b["rgb"].view(np.uint32))


def test_batch_mixes_transformed_and_world_scenes():
    from eos_amd import _lib as E
    print("transform batch", flush=True)
    ss = [CASES["inst8_rot"](), scenes.config_a(64, 48), CASES["inst8_axis"]()]
    gs = [E.Scene(s, leaf_max=4) for s in ss]
    singles = [g.render()["rgba"] for g in gs]
    for use_graph in (False, True):
        outs, _ = E.render_batch(gs, depth=2, use_graph=use_graph)
        for k, (s, o) in enumerate(zip(ss, outs)):
            want = gs[k].render(depth=2)["rgba"]
            assert np.array_equal(o, want), (k, use_graph)
    assert singles[0].shape == (72, 96, 4)


def test_set_transforms_rules():
    from eos_amd import _lib as E
    s = CASES["inst8_rot"]()
    print("transform api rules", flush=True)
    g = E.Scene(s, leaf_max=4)
    base = g.render(want_hits=True)
    # a new transformation set needs a rebuild
    g.set_transforms(s.xforms, s.xform_index)
    with pytest.raises(E.EosError) as ei:
        g.render()
    assert ei.value.status == E.EOS_ERR_STATE
    g.build(4)
    again = g.render(want_hits=True)
    assert np.array_equal(again["hit"], base["hit"]) and np.array_equal(again["rgba"], base["rgba"])
    # singular matrix, index out of range
    bad = s.xforms.copy()
    bad[3] = 0.0
    with pytest.raises(E.EosError) as ei:
        g.set_transforms(bad, s.xform_index)
    assert ei.value.status == E.EOS_ERR_INVALID
    idx = s.xform_index.copy()
    idx[7] = len(s.xforms)
    with pytest.raises(E.EosError) as ei:
        g.set_transforms(s.xforms, idx)
    assert ei.value.status == E.EOS_ERR_INVALID
    # identity matrices everywhere: the world-space frame
    w = scenes.SceneData(s.verts, s.normals, s.mat, s.materials, s.lights, s.camera, s.depth, "plain")
    gw = E.Scene(w, leaf_max=4)
    ident = np.tile(np.eye(4, dtype=np.float32).ravel(), (len(s.xforms), 1))
    g.set_transforms(ident, s.xform_index)
    g.build(4)
    a, b = g.render(want_hits=True), gw.render(want_hits=True)
    assert np.array_equal(a["hit"], b["hit"]) and np.array_equal(a["rgba"], b["rgba"])


def test_adjoint_with_transforms_matches_fp64_forward_mode():
    """Gradients on a transformed scene: hit records come from the
    transformed path, the FP64 forward mode from the oracle's records."""
    from eos_amd import _lib as E
    s = CASES["inst8_rot"]()
    print("transform adjoint", flush=True)
    g = E.Scene(s, leaf_max=4)
    tm, tl = scenes.perturb_params(s, 99)
    g.set_materials(tm)
    g.set_lights(tl)
    target = g.render(want_rgb=True)["rgb"]
    g.set_materials(s.materials)
    g.set_lights(s.lights)
    fwd = g.render(want_rgb=True, want_hits=True, keep_records=True)
    dl = 2.0 * (fwd["rgb"] - target)
    gm, gl, _ = g.adjoint(dl)
    gm2, gl2, _ = g.adjoint(dl, recompute=True)
    close = lambda a, b: np.allclose(a, b, rtol=1e-5, atol=1e-6 * np.abs(b).max())  # noqa: E731
    assert close(gm, gm2) and close(gl, gl2)          # FP32 workgroup sums: order differs
    out = po.Oracle(s).render(want_recs=True)
    assert (out["hit"] == fwd["hit"]).mean() >= 0.9999
    prm = po.params_vector(s.materials, s.lights)
    ref = po.grad_fwd64(out["recs"], s.camera["position"], len(s.materials), len(s.lights), prm, target)
    g_gpu = np.concatenate([gm.ravel(), gl.ravel()])
    scale = np.abs(ref).max()
    err = np.abs(g_gpu - ref) / np.maximum(np.abs(ref), 1e-3 * scale)
    print(f"  max rel err vs forward mode {err.max():.2e}", flush=True)
    assert err.max() <= 1e-3


def test_full_size_instanced_frame_sampled_rows():
    """Config F at its full size (4,194,304 instanced triangles, 3840x2160): the
    GPU build bit-exact against the restatement and the primary hits of every
    32nd row equal to the restatement's walk of the same rays."""
    from eos_amd import _lib as E
    s = scenes.config_f()
    print(f"transform full size n={s.n_tri}", flush=True)
    g = E.Scene(s, leaf_max=4)
    b = po.lbvh_build(s.verts, 4, s.xforms, s.xform_index)
    assert np.array_equal(g.debug_read("perm"), b["perm"])
    assert np.array_equal(_u32(g.debug_read("nodes4"))[:, :28], _u32(b["nodes4"])[:, :28])
    go = g.render(want_hits=True)
    W, H = s.width, s.height
    lib = po.lib()
    cam = np.array(s.camera, copy=True)
    rows = np.arange(0, H, 32)
    dirs = np.zeros((len(rows), W, 3), np.float32)
    for i, y in enumerate(rows):
        for x in range(W):
            lib.orc_camera_ray(cam.ctypes.data, x, int(y), dirs[i, x].ctypes.data)
    dirs = dirs.reshape(-1, 3)
    orig = np.tile(np.asarray(s.camera["position"], np.float32).reshape(1, 3), (len(dirs), 1))
    r = po.lbvh_trace(b, orig, dirs, width=4, xforms=s.xforms)
    want = r["hit"].reshape(len(rows), W)          # original triangle indices
    got = go["hit"][rows]
    agree = float((got == want).mean())
    print(f"  sampled rows agree={agree} hit fraction={(got >= 0).mean():.3f}", flush=True)
    assert agree >= 0.9999
    assert (got >= 0).mean() > 0.5
