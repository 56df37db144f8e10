"""Primary tile test (eos_launch_config.cull): tiles whose rays all miss the
root node's child boxes are resolved as misses before the primary launch, and
only the other tiles are traced.  Frames, hit IDs, float colours, ray counts,
records and gradients are identical with and without it, for cameras that see
the scene, look away from it, sit inside its box or graze it, for row bands,
anti-aliased frames, ray trees, transformed triangles and frames in flight.
The geometry of the test (no tile holding a ray that meets a child box is
resolved) is checked on the CPU in test_cull_geometry.py, against float64 slab
tests of every pixel's ray, and here through the frames it leaves unchanged."""
import numpy as np
import pytest

from eos_amd import scenes

pytestmark = pytest.mark.gpu


def _same(a, b, what):
    assert np.array_equal(a["rgba"], b["rgba"]), what
    if a.get("hit") is not None:
        assert np.array_equal(a["hit"], b["hit"]), what
    if a.get("rgb") is not None:
        assert np.array_equal(a["rgb"].view(np.uint32), b["rgb"].view(np.uint32)), what
    for k in ("primary_rays", "shadow_rays", "secondary_rays", "hits"):
        assert a["stats"][k] == b["stats"][k], (what, k)


def _cameras(s):
    W, H = 320, 184   # 184 rows: the last tile row is partial
    return {
        "scene camera": np.array(s.camera, copy=True),
        "small frame": scenes.make_camera((0.0, 0.4, 3.3), (0.0, 0.0, 0.0), 60.0, W, H),
        "looking away": scenes.make_camera((0.0, 0.4, 3.3), (0.0, 0.0, 10.0), 60.0, W, H),
        "inside the box": scenes.make_camera((0.0, 0.0, 0.3), (0.0, 0.0, 2.0), 60.0, W, H),
        "grazing": scenes.make_camera((1.2, 0.0, 0.0), (1.2, 0.0, 5.0), 60.0, W, H),
        "far": scenes.make_camera((0.0, 0.0, 40.0), (0.0, 0.0, 0.0), 60.0, W, H),
    }


def test_tile_test_leaves_frames_unchanged():
    from eos_amd import _lib as E
    s = scenes.config_c(width=640, height=360, nu=256, nv=128)
    g = E.Scene(s, leaf_max=4)
    for name, cam in _cameras(s).items():
        print(f"tile test: camera {name}", flush=True)
        g.set_launch_config(cull=0)
        a = g.render(cam, want_rgb=True, want_hits=True)
        assert a["stats"]["primary_culled"] == 0
        g.set_launch_config(cull=1)
        b = g.render(cam, want_rgb=True, want_hits=True)
        _same(a, b, name)
        culled = b["stats"]["primary_culled"]
        assert 0 <= culled <= b["stats"]["primary_rays"]
        if name == "looking away":
            assert culled == b["stats"]["primary_rays"], culled
        if name in ("scene camera", "small frame"):
            assert culled > b["stats"]["primary_rays"] // 4, culled
        if name == "inside the box":
            assert culled == 0
        # counting frames never cull: their visit counts restate every ray's walk
        c = g.render(cam, count_visits=True)
        assert c["stats"]["primary_culled"] == 0
    g.set_launch_config()


def test_tile_test_bands_aa_and_frames_in_flight():
    import torch
    from eos_amd import _lib as E
    s = scenes.config_c(width=640, height=360, nu=256, nv=128)
    g = E.Scene(s, leaf_max=4)
    for cull in (0, 1):
        g.set_launch_config(cull=cull)
        if cull == 0:
            ref_rows = [g.render_rows(r0, r1, want_rgb=True, want_hits=True) for r0, r1 in ((0, 123), (123, 360))]
            ref_aa = g.render(aa_samples=2, want_rgb=True)
        else:
            for (r0, r1), ref in zip(((0, 123), (123, 360)), ref_rows):
                print(f"tile test: rows {r0}..{r1}", flush=True)
                _same(ref, g.render_rows(r0, r1, want_rgb=True, want_hits=True), f"rows {r0}..{r1}")
            print("tile test: 2x2 anti-aliasing", flush=True)
            _same(ref_aa, g.render(aa_samples=2, want_rgb=True), "aa")
    # frames in flight, cameras alternating between a culled and an unculled view
    cams = np.stack([np.array(c, copy=True) for c in list(_cameras(s).values())[1:4]] * 2)
    W, H = int(cams[0]["width"]), int(cams[0]["height"])
    outs = {}
    for cull in (0, 1):
        g.set_launch_config(cull=cull)
        bufs = [torch.zeros(W * H * 4, dtype=torch.uint8, device="cuda:0") for _ in range(len(cams))]
        st = g.render_frames(cams, [b.data_ptr() for b in bufs])
        torch.cuda.synchronize()
        outs[cull] = ([b.cpu().numpy() for b in bufs], st)
    print("tile test: frames in flight", flush=True)
    for i, (x, y) in enumerate(zip(outs[0][0], outs[1][0])):
        assert np.array_equal(x, y), i
        for k in ("shadow_rays", "secondary_rays", "hits"):
            assert outs[0][1][i][k] == outs[1][1][i][k], (i, k)
    g.set_launch_config()


def test_tile_test_ray_trees_transforms_and_gradients():
    from eos_amd import _lib as E
    cases = {
        "glass (ray trees)": scenes.glass(320, 184, level=2, depth=3),
        "instanced (transformations)": scenes.config_f(320, 184, instances=9, nu=32, nv=16),
    }
    for name, s in cases.items():
        print(f"tile test: {name}", flush=True)
        g = E.Scene(s, leaf_max=4)
        g.set_launch_config(cull=0)
        a = g.render(want_rgb=True, want_hits=True)
        g.set_launch_config(cull=1)
        b = g.render(want_rgb=True, want_hits=True)
        _same(a, b, name)
    # adjoint records of a frame with the tile test
    s = scenes.config_e(width=96, height=96, nu=100, nv=50, n_materials=16, n_lights=4)
    g = E.Scene(s, leaf_max=4)
    dl = np.random.default_rng(3).normal(size=(96, 96, 3)).astype(np.float32)
    grads = {}
    for cull in (0, 1):
        g.set_launch_config(cull=cull)
        g.render(keep_records=True)
        grads[cull] = g.adjoint(dl)
        grads[(cull, "recompute")] = g.adjoint(dl, recompute=True)
    print("tile test: adjoint (replayed and recomputed records)", flush=True)
    for k in (1, (0, "recompute"), (1, "recompute")):
        assert np.array_equal(grads[0][0], grads[k][0]) and np.array_equal(grads[0][1], grads[k][1]), k
        assert grads[0][2]["records"] == grads[k][2]["records"], k


def test_tile_test_frame_of_over_a_million_tiles():
    """8448 x 8200 pixels: 1,082,400 tiles, more workgroups of the tile test
    than the compaction sums directly, so the separate scan of their counts
    runs (k_cull_scan)."""
    from eos_amd import _lib as E
    s = scenes.config_c(width=320, height=184, nu=64, nv=32)
    g = E.Scene(s, leaf_max=4)
    cam = scenes.make_camera((0.0, 0.4, 3.3), (0.0, 0.0, 0.0), 60.0, 8448, 8200)
    print("tile test: 8448 x 8200 frame", flush=True)
    g.set_launch_config(cull=0)
    a = g.render(cam, depth=1, want_hits=True)
    g.set_launch_config(cull=1)
    b = g.render(cam, depth=1, want_hits=True)
    _same(a, b, "8448 x 8200")
    assert b["stats"]["primary_culled"] > b["stats"]["primary_rays"] // 4
    g.set_launch_config()
