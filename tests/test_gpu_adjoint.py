"""HIP adjoint vs FP64 CPU references (hit sets held fixed).

Bar: relative error <= 1e-3 per parameter, where the relative error of a
parameter is |g - g_ref| / max(|g_ref|, 1e-3 * max_j |g_ref_j|) (the floor keeps
parameters whose gradient is ~0 from dividing by zero).
"""
import numpy as np
import pytest

from eos_amd import scenes
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

REL_TOL = 1e-3


def _rel_err(g, ref):
    scale = np.abs(ref).max()
    return np.abs(g - ref) / np.maximum(np.abs(ref), 1e-3 * scale)


def _setup(width=96, n_materials=16, n_lights=8):
    from eos_amd import _lib as E
    s = scenes.config_e(width, width, nu=100, nv=50, n_materials=n_materials, n_lights=n_lights)
    g = E.Scene(s, leaf_max=4)
    tm, tl = scenes.perturb_params(s, 1234)
    g.set_materials(tm)
    g.set_lights(tl)
    target = g.render(want_rgb=True)["rgb"]
    g.set_materials(s.materials)
    g.set_lights(s.lights)
    return s, g, target


def test_adjoint_matches_fp64_forward_mode_and_central_differences():
    s, g, target = _setup()
    print(f"adjoint: {s.n_tri} tris {s.width}x{s.height} M={len(s.materials)} L={len(s.lights)}", flush=True)
    fwd = g.render(want_rgb=True, want_hits=True, keep_records=True)
    dl = 2.0 * (fwd["rgb"] - target)
    gm, gl, st = g.adjoint(dl)
    g_gpu = np.concatenate([gm.ravel(), gl.ravel()])
    assert st["records"] > 0 and st["retrace_ms"] == 0.0          # replayed, not re-traced
    orc = po.Oracle(s)
    out = orc.render(want_recs=True)
    assert (out["hit"] == fwd["hit"]).mean() >= 0.9999
    prm = po.params_vector(s.materials, s.lights)
    ref = po.grad_fwd64(out["recs"], s.camera["position"], len(s.materials), len(s.lights), prm, target)
    err = _rel_err(g_gpu, ref)
    print(f"  max rel err vs forward mode {err.max():.2e}", flush=True)
    assert err.max() <= REL_TOL
    rng = np.random.default_rng(5)
    which = np.sort(rng.choice(len(prm), 64, replace=False))
    fd = po.grad_fd64(out["recs"], s.camera["position"], len(s.materials), len(s.lights), prm, target, which,
                      h=1e-3)
    scale = np.abs(ref).max()
    e2 = np.abs(g_gpu[which] - fd) / np.maximum(np.abs(fd), 1e-3 * scale)
    print(f"  max rel err vs central differences (64 params, h=1e-3) {e2.max():.2e}", flush=True)
    assert e2.max() <= REL_TOL


def test_adjoint_replay_equals_recompute_and_accumulates():
    s, g, target = _setup(64, 8, 4)
    fwd = g.render(want_rgb=True, keep_records=True)
    dl = 2.0 * (fwd["rgb"] - target)
    gm1, gl1, st1 = g.adjoint(dl)
    gm2, gl2, st2 = g.adjoint(dl, recompute=True)
    assert st2["retrace_ms"] > 0.0
    # the same records in another order (slots are claimed by atomics): the sums
    # are exact integer sums, so the gradients are the same bits
    assert np.array_equal(gm1, gm2) and np.array_equal(gl1, gl2)
    # records are dropped when parameters change: the next adjoint re-traces
    g.set_lights(s.lights)
    gm3, gl3, st3 = g.adjoint(dl)
    assert st3["retrace_ms"] > 0.0 and np.array_equal(gm1, gm3) and np.array_equal(gl1, gl3)
    for _ in range(3):
        gm4, gl4, _ = g.adjoint(dl, recompute=True)
        assert np.array_equal(gm1, gm4) and np.array_equal(gl1, gl4)
    close = lambda a, b: np.allclose(a, b, rtol=1e-12, atol=0.0)  # noqa: E731
    # gradients accumulate into the caller's arrays (device path)
    import torch
    gm_d = torch.zeros((len(s.materials), 7), dtype=torch.float64, device="cuda:0")
    gl_d = torch.zeros((len(s.lights), 6), dtype=torch.float64, device="cuda:0")
    dl_d = torch.tensor(dl, device="cuda:0")
    torch.cuda.synchronize()
    for _ in range(2):
        g.adjoint_device(s.camera, s.depth, dl_d.data_ptr(), gm_d.data_ptr(), gl_d.data_ptr())
    assert close(gm_d.cpu().numpy(), 2 * gm1) and close(gl_d.cpu().numpy(), 2 * gl1)


def test_autograd_function():
    import torch
    from eos_amd import autograd as A
    s, g, target = _setup(48, 6, 3)
    params = [p.requires_grad_(True) for p in A.params_of(s)]
    img = A.render(g, *params)
    tgt = torch.tensor(target, device=img.device)
    loss = ((img - tgt) ** 2).sum()
    loss.backward()
    got = np.concatenate([params[0].grad.cpu().numpy().ravel(), params[1].grad.cpu().numpy().ravel(),
                          params[2].grad.cpu().numpy().ravel()])
    gm, gl, _ = g.adjoint(2.0 * (img.detach().cpu().numpy() - target))
    want = np.concatenate([gm[:, 0:3].ravel(), gm[:, 3:6].ravel(), gm[:, 6]])
    assert np.allclose(got, want, rtol=1e-5, atol=1e-6)
    assert np.allclose(params[3].grad.cpu().numpy(), gl[:, 0:3], rtol=1e-5, atol=1e-6)
    assert np.allclose(params[4].grad.cpu().numpy(), gl[:, 3:6], rtol=1e-5, atol=1e-6)


def test_autograd_two_renders_in_one_loss():
    """Two forwards on one scene with different parameters before one backward:
    each call's gradients are its own (the second render replaced the scene's
    records; the first call's backward re-traces with its own parameters),
    checked against the oracle's FP64 forward mode for each parameter set."""
    import torch
    from eos_amd import autograd as A
    s, g, target = _setup(48, 6, 3)
    p1 = [p.requires_grad_(True) for p in A.params_of(s)]
    tm, tl = scenes.perturb_params(s, 77, rel=0.05)
    s2 = scenes.SceneData(s.verts, s.normals, s.mat, tm, tl, s.camera, s.depth)
    p2 = [p.requires_grad_(True) for p in A.params_of(s2)]
    tgt = torch.tensor(target, device="cuda:0")
    img1 = A.render(g, *p1)
    img2 = A.render(g, *p2)
    loss = ((img1 - tgt) ** 2).sum() + 0.5 * ((img2 - tgt) ** 2).sum()
    loss.backward()
    orc = po.Oracle(s)
    for params, data, w, img in ((p1, s, 1.0, img1), (p2, s2, 0.5, img2)):
        orc.set_params(data.materials, data.lights)
        recs = orc.render(want_recs=True)["recs"]
        prm = po.params_vector(data.materials, data.lights)
        ref = w * po.grad_fwd64(recs, s.camera["position"], len(data.materials), len(data.lights), prm, target)
        got = np.concatenate([torch.cat([params[0].grad, params[1].grad, params[2].grad[:, None]], 1)
                              .cpu().numpy().ravel(),
                              torch.cat([params[3].grad, params[4].grad], 1).cpu().numpy().ravel()])
        err = _rel_err(got, ref)
        print(f"  two renders: max rel err {err.max():.2e}", flush=True)
        assert err.max() <= REL_TOL
    # the scene keeps the parameters of the last forward
    assert np.array_equal(g._mats["kd"][:, :3], s2.materials["kd"][:, :3])


def test_adjoint_config_e_full_size():
    """Config E at its stated size: 1,000,000 triangles, 256 materials, 8 lights,
    1024 x 1024, depth 2, 1,840 parameters.  Upstream gradient on every 16th row
    (the oracle's share); GPU gradients against the FP64 forward mode of those
    rows and against FP64 central differences (h = 1e-3) on 64 parameters."""
    from eos_amd import _lib as E
    s = scenes.config_e()
    print(f"adjoint full size: {s.n_tri} tris {s.width}x{s.height} M={len(s.materials)} L={len(s.lights)}",
          flush=True)
    g = E.Scene(s, leaf_max=4)
    tm, tl = scenes.perturb_params(s, 1234)
    g.set_materials(tm)
    g.set_lights(tl)
    target = g.render(want_rgb=True)["rgb"]
    g.set_materials(s.materials)
    g.set_lights(s.lights)
    fwd = g.render(want_rgb=True, want_hits=True, keep_records=True)
    rows = np.zeros(s.height, bool)
    rows[::16] = True
    dl = np.where(rows[:, None, None], 2.0 * (fwd["rgb"] - target), 0.0).astype(np.float32)
    gm, gl, st = g.adjoint(dl)
    assert st["retrace_ms"] == 0.0
    g_gpu = np.concatenate([gm.ravel(), gl.ravel()])
    assert g_gpu.size == 1840
    out = po.Oracle(s).render(row_start=0, row_step=16, want_recs=True)
    assert (out["hit"][::16] == fwd["hit"][::16]).mean() >= 0.9999
    recs = out["recs"]                  # rows the oracle did not render hold no hits
    tgt = target
    prm = po.params_vector(s.materials, s.lights)
    ref = po.grad_fwd64(recs, s.camera["position"], len(s.materials), len(s.lights), prm, tgt)
    err = _rel_err(g_gpu, ref)
    print(f"  max rel err vs forward mode {err.max():.2e}", flush=True)
    assert err.max() <= REL_TOL
    rng = np.random.default_rng(9)
    which = np.sort(rng.choice(len(prm), 64, replace=False))
    fd = po.grad_fd64(recs, s.camera["position"], len(s.materials), len(s.lights), prm, tgt, which, h=1e-3)
    e2 = np.abs(g_gpu[which] - fd) / np.maximum(np.abs(fd), 1e-3 * np.abs(ref).max())
    print(f"  max rel err vs central differences {e2.max():.2e}", flush=True)
    assert e2.max() <= REL_TOL
