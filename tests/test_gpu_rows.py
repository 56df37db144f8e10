"""eos_render_rows: a frame split into bands of rows (one GPU's share of the
reference's split-frame renderer) stitches back to the whole frame, byte for
byte, whatever the band boundaries (tile-aligned or not)."""
import numpy as np
import pytest

from eos_amd import scenes

pytestmark = pytest.mark.gpu

CASES = {
    "soup": lambda: scenes.config_b(96, 70, n=20000),
    "glass": lambda: scenes.glass(64, 45, level=2),
    "instanced": lambda: scenes.config_f(72, 51, instances=8, nu=16, nv=8),
}


@pytest.mark.parametrize("aa", [0, 2])
@pytest.mark.parametrize("name", sorted(CASES))
def test_bands_stitch_to_the_frame(name, aa):
    from eos_amd import _lib as E
    s = CASES[name]()
    H = s.height
    print(f"rows case={name} aa={aa} {s.width}x{H}", flush=True)
    g = E.Scene(s, leaf_max=4)
    full = g.render(want_rgb=True, want_hits=True, aa_samples=aa)
    for cuts in ([0, H], [0, H // 2, H], [0, 5, 13, 24, H], [0, 8, 16, H], [0, H // 3, 2 * (H // 3)]):
        rgba = np.concatenate([g.render_rows(a, b, aa_samples=aa)["rgba"] for a, b in zip(cuts, cuts[1:])])
        rgb = np.concatenate([g.render_rows(a, b, want_rgb=True, aa_samples=aa)["rgb"] for a, b in zip(cuts, cuts[1:])])
        hit = np.concatenate([g.render_rows(a, b, want_hits=True, aa_samples=aa)["hit"] for a, b in zip(cuts, cuts[1:])])
        rows = cuts[-1]
        assert np.array_equal(rgba, full["rgba"][:rows]), cuts
        assert np.array_equal(rgb.view(np.uint32), full["rgb"][:rows].view(np.uint32)), cuts
        assert np.array_equal(hit, full["hit"][:rows]), cuts
    st = g.render_rows(10, 20)["stats"]
    assert st["primary_rays"] == 10 * s.width


def test_rows_rules():
    from eos_amd import _lib as E
    s = scenes.config_a(32, 24)
    g = E.Scene(s, leaf_max=4)
    print("rows rules", flush=True)
    for a, b in ((0, 0), (5, 3), (-1, 4), (0, 25)):
        with pytest.raises(E.EosError) as ei:
            g.render_rows(a, b)
        assert ei.value.status == E.EOS_ERR_INVALID
