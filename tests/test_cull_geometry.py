"""CPU check of the primary tile test's geometry (k_cull_tiles in
eos_amd/csrc/render.hip, restated here in numpy float64): no 8x8 tile that
the test resolves holds a camera ray -- computed with camera_dir's float32
operations, with and without an anti-aliasing offset -- whose float64 slab test
meets a child box of the root node (boxes widened by 1e-6 relative, beyond the
GPU slab test's rounding).  Cameras: the scene's own, row bands, inside the
scene box, looking away, grazing, far, and the hit-dense view.  The trees are
the oracle's LBVH restatement, equal to the GPU's bit for bit."""
import numpy as np
import pytest

from eos_amd import scenes
from oracle import pyoracle as O

EMPTY = 0x7FFFFFFE


def _root_boxes(b):
    nd = b["nodes4"][b["root4"]]
    lo = np.stack([nd[0:4], nd[8:12], nd[16:20]], 1).astype(np.float64)
    hi = np.stack([nd[4:8], nd[12:16], nd[20:24]], 1).astype(np.float64)
    ref = nd[24:28].view(np.int32)
    return [(lo[k], hi[k]) for k in range(4) if ref[k] != EMPTY]


def _traced_tiles(boxes, cam, row0, band_h):
    """k_cull_tiles: 1 where the tile is traced."""
    W, H = int(cam["width"]), int(cam["height"])
    r, u, v = (np.array(cam[k][:3], np.float64) for k in ("right", "up", "view"))
    Op = np.array(cam["position"][:3], np.float64)
    nw, nh = W / 2.0, H / 2.0
    ty, tx = np.meshgrid(np.arange((band_h + 7) // 8), np.arange((W + 7) // 8), indexing="ij")
    x0, x1 = tx * 8.0 - 1, tx * 8.0 + 9
    y0, y1 = ty * 8.0 + row0 - 1, ty * 8.0 + row0 + 9

    def d(px, py):
        return ((px - nw) / nw)[..., None] * r + ((py - nh) / nh)[..., None] * u + v

    D = [d(x0, y0), d(x1, y0), d(x1, y1), d(x0, y1)]
    C = d(0.5 * (x0 + x1), 0.5 * (y0 + y1))
    N = []
    for k in range(4):
        n = np.cross(D[k], D[(k + 1) % 4])
        N.append(n * np.where(np.sum(n * C, -1) < 0, -1.0, 1.0)[..., None])
    front = np.all([np.sum(C * Dk, -1) > 0 for Dk in D], 0)
    N.append(np.where(front[..., None], C, 0.0))
    missed = np.ones(tx.shape, bool)
    for L, Hh in boxes:
        scale = np.abs(Op).sum() + np.sum(np.maximum(np.abs(L - Op), np.abs(Hh - Op)) + np.maximum(np.abs(L), np.abs(Hh)))
        outside = np.zeros(tx.shape, bool)
        for n in N:
            far = np.sum(n * (np.where(n >= 0, Hh, L) - Op), -1)
            outside |= far < -1e-3 * np.sqrt(np.sum(n * n, -1)) * scale
        missed &= outside
    return ~missed


def _rays_meeting_boxes(boxes, cam, row0, band_h, ox, oy):
    W, H = int(cam["width"]), int(cam["height"])
    nw, nh = np.float32(W / 2.0), np.float32(H / 2.0)
    inw, inh = np.float32(1) / nw, np.float32(1) / nh
    ys, xs = np.meshgrid(np.arange(band_h) + row0, np.arange(W), indexing="ij")
    a = ((xs.astype(np.float32) + np.float32(ox)) - nw) * inw
    bb = ((ys.astype(np.float32) + np.float32(oy)) - nh) * inh
    r, u, v = (np.array(cam[k][:3], np.float32) for k in ("right", "up", "view"))
    dd = a[..., None] * r + bb[..., None] * u + v
    dd = (dd / np.linalg.norm(dd, axis=-1, keepdims=True)).astype(np.float64)
    o = np.array(cam["position"][:3], np.float64)
    with np.errstate(divide="ignore"):
        inv = 1.0 / dd
    meet = np.zeros(xs.shape, bool)
    for L, Hh in boxes:
        ext = np.abs(Hh - L).max() * 1e-6 + 1e-6 * np.abs(o).max()
        with np.errstate(invalid="ignore"):
            t0, t1 = (L - ext - o) * inv, (Hh + ext - o) * inv
            tn = np.nanmax(np.minimum(t0, t1), -1)
            tf = np.nanmin(np.maximum(t0, t1), -1)
        meet |= np.maximum(tn, 0) <= tf
    return meet, ys - row0, xs


def _check(b, cam, row0=0, band_h=None, ox=0.0, oy=0.0):
    H = int(cam["height"])
    band_h = H - row0 if band_h is None else band_h
    boxes = _root_boxes(b)
    traced = _traced_tiles(boxes, cam, row0, band_h)
    meet, yb, xs = _rays_meeting_boxes(boxes, cam, row0, band_h, ox, oy)
    tile = (yb // 8) * traced.shape[1] + xs // 8
    assert not np.any(meet & ~traced.ravel()[tile])
    return float(traced.mean())


@pytest.fixture(scope="module")
def sphere():
    s = scenes.config_c(width=640, height=360, nu=128, nv=64)
    return s, O.lbvh_build(s.verts.reshape(-1, 3) if s.verts.ndim == 1 else s.verts, leaf_max=4)


def test_scene_camera_bands_and_aa(sphere):
    s, b = sphere
    frac = _check(b, s.camera)
    assert frac < 0.6   # most tiles of this view are resolved
    _check(b, s.camera, ox=0.999, oy=0.999)
    _check(b, s.camera, row0=100, band_h=77)


@pytest.mark.parametrize("eye,target,expect", [
    ((0.0, 0.0, 0.3), (0.0, 0.0, 2.0), "all"),      # inside the scene box
    ((0.0, 0.4, 3.3), (0.0, 0.0, 10.0), "none"),    # looking away
    ((1.2, 0.0, 0.0), (1.2, 0.0, 5.0), None),       # grazing the box
    ((0.0, 0.0, 40.0), (0.0, 0.0, 0.0), None),      # far away
    ((0.3, 1.6, 0.2), (0.0, 0.0, 0.0), "all"),      # close above
])
def test_other_cameras(sphere, eye, target, expect):
    s, b = sphere
    frac = _check(b, scenes.make_camera(eye, target, 60.0, 320, 184))
    if expect == "all":
        assert frac == 1.0
    if expect == "none":
        assert frac == 0.0


def test_hit_dense_view():
    d = scenes.config_c_dense(width=480, height=272, nu=128, nv=64)
    b = O.lbvh_build(d.verts.reshape(-1, 3) if d.verts.ndim == 1 else d.verts, leaf_max=4)
    assert _check(b, d.camera) == 1.0
