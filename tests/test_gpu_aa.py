"""Anti-aliased frames (eos_render with aa_samples = n) vs the oracle's
restatement of RayTracer::renderWithAA, which reproduces the reference's own
anti-aliased images byte for byte (tests/test_golden.py).

Bars: RGBA within 1/255 on >= 99.99% of pixels (a pixel mixes n*n rays, so a
grazing-edge hit difference in one sample moves the pixel by a fraction of a
level), float colour within 2e-5 on the rest, ray counts exact.
"""
import numpy as np
import pytest

from eos_amd import scenes
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

CASES = {
    "config_a_64x48_n2": (lambda: scenes.config_a(64, 48), 2),
    "config_a_37x23_n4": (lambda: scenes.config_a(37, 23), 4),
    "sphere_16k_n3": (lambda: scenes.config_c(80, 45, nu=128, nv=64), 3),
    "instanced_n2": (lambda: scenes.config_f(64, 48, instances=8, nu=20, nv=10), 2),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_aa_matches_oracle(name):
    from eos_amd import _lib as E
    make, n = CASES[name]
    s = make()
    print(f"aa case={name} n={n} {s.width}x{s.height} depth={s.depth}", flush=True)
    g = E.Scene(s, leaf_max=4)
    go = g.render(want_rgb=True, aa_samples=n)
    oo = po.Oracle(s).render_aa(n)
    d = np.abs(go["rgba"].astype(int) - oo["rgba"].astype(int)).max(-1)
    rd = np.abs(go["rgb"] - oo["rgb"]).max(-1)
    print(f"  exact={float((d == 0).mean()):.6f} within1={float((d <= 1).mean()):.6f} "
          f"rgb<=2e-5={float((rd <= 2e-5).mean()):.6f}", flush=True)
    assert (d <= 1).mean() >= 0.9999
    assert (rd <= 2e-5).mean() >= 0.999
    st, oc = go["stats"], oo["counts"]
    assert st["primary_rays"] == oc["primary"] == n * n * s.width * s.height
    assert st["shadow_rays"] == oc["shadow"] and st["secondary_rays"] == oc["secondary"]
    assert go["rgba"][..., 3].min() == 255


def test_one_sample_is_the_plain_frame_and_hits_are_sample_zero():
    from eos_amd import _lib as E
    s = scenes.config_b(96, 64, n=20000)
    print("aa n=1 and hit_out", flush=True)
    g = E.Scene(s, leaf_max=4)
    plain = g.render(want_rgb=True, want_hits=True)
    one = g.render(want_rgb=True, want_hits=True, aa_samples=1)
    assert np.array_equal(plain["rgba"], one["rgba"])
    assert np.array_equal(plain["rgb"].view(np.uint32), one["rgb"].view(np.uint32))
    four = g.render(want_rgb=True, want_hits=True, aa_samples=4)
    assert np.array_equal(four["hit"], plain["hit"])       # sample (0, 0) is the pixel-corner ray
    assert not np.array_equal(four["rgba"], plain["rgba"])


def test_aa_rules():
    from eos_amd import _lib as E
    s = scenes.config_a(16, 16)
    g = E.Scene(s, leaf_max=4)
    print("aa rules", flush=True)
    with pytest.raises(E.EosError) as ei:
        g.render(aa_samples=2, keep_records=True)
    assert ei.value.status == E.EOS_ERR_INVALID
    with pytest.raises(E.EosError):
        g.render(aa_samples=17)
    # the device-output route gives the host route's image
    import torch
    out = torch.empty(16 * 16 * 4, dtype=torch.uint8, device="cuda:0")
    g.render_device(s.camera, s.depth, out.data_ptr(), aa_samples=3)
    host = g.render(aa_samples=3)["rgba"]
    assert np.array_equal(out.cpu().numpy().reshape(16, 16, 4), host)
