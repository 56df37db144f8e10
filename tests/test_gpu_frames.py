"""eos_render_frames (two frames in flight) against eos_render frame by frame:
same bytes, same ray counts, on chains, ray trees and transformed scenes, with
cameras of different sizes and with outputs that alias."""
import numpy as np
import pytest

from eos_amd import scenes

pytestmark = pytest.mark.gpu


def _cams(s):
    a = np.array(s.camera, copy=True)
    b = scenes.make_camera((0.6, 0.7, 2.4), (0.5, 0.45, 0.5), 50.0, 72, 40)
    c = s.with_camera(33, 57).camera
    d = np.array(s.camera, copy=True)
    d["position"] = d["position"] * np.float32(1.15)
    return np.stack([a, b, c, d, a])


CASES = {
    "soup": lambda: scenes.config_b(64, 48, n=20000),
    "glass": lambda: scenes.glass(64, 48, level=2),
    "instanced": lambda: scenes.config_f(64, 48, instances=8, nu=16, nv=8),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_frames_equal_single_renders(name):
    import torch
    from eos_amd import _lib as E
    s = CASES[name]()
    cams = _cams(s)
    print(f"frames case={name} n={len(cams)}", flush=True)
    g = E.Scene(s, leaf_max=4)
    outs = [torch.zeros(int(c["width"]) * int(c["height"]) * 4, dtype=torch.uint8, device="cuda:0") for c in cams]
    st = g.render_frames(cams, [o.data_ptr() for o in outs])
    for i, c in enumerate(cams):
        one = g.render(camera=c)
        got = outs[i].cpu().numpy().reshape(int(c["height"]), int(c["width"]), 4)
        assert np.array_equal(got, one["rgba"]), i
        for k in ("primary_rays", "shadow_rays", "secondary_rays", "hits"):
            assert st[i][k] == one["stats"][k], (i, k)
        assert st[i]["render_ms"] > 0.0
    # counting frames: the walk counters per frame equal the single frame's
    st2 = g.render_frames(cams[:2], [o.data_ptr() for o in outs[:2]], count_visits=True)
    for i in range(2):
        one = g.render(camera=cams[i], count_visits=True)["stats"]
        assert st2[i]["node_visits"] == one["node_visits"] and st2[i]["tri_tests"] == one["tri_tests"]


def test_aliased_output_holds_the_last_frame_and_rules():
    import torch
    from eos_amd import _lib as E
    s = scenes.config_a(48, 40)
    print("frames aliasing", flush=True)
    g = E.Scene(s, leaf_max=4)
    cams = np.stack([np.array(s.camera, copy=True) for _ in range(6)])
    for i in range(6):
        cams[i]["position"] = cams[i]["position"] + np.float32([0.1 * i, 0.0, 0.0])
    out = torch.zeros(48 * 40 * 4, dtype=torch.uint8, device="cuda:0")
    g.render_frames(cams, [out.data_ptr()] * 6)
    last = g.render(camera=cams[-1])["rgba"]
    assert np.array_equal(out.cpu().numpy().reshape(40, 48, 4), last)
    assert g.render_frames(cams[:0], []) == []
    L = E.lib()
    import ctypes as C
    o = E.RenderOpts(1, E.EOS_MEM_HOST, None, None, 0, 0, 0)
    ptrs = (C.c_void_p * 1)(out.data_ptr())
    assert L.eos_render_frames(g.h, cams.ctypes.data, 1, C.byref(o), ptrs, None) == E.EOS_ERR_INVALID
