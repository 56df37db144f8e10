"""eos_trace_rays_static, the allocation-free ray query, against eos_trace_rays
through the C ABI: two calls in a row on each tree width and query mode, the
second with fewer rays, with bit-identical results; the size limit and the
device-memory requirement are refused with clear errors."""
import ctypes as C

import numpy as np
import pytest

from eos_amd import scenes

pytestmark = pytest.mark.gpu


def _rays(s, n, seed):
    rng = np.random.default_rng(seed)
    v = s.verts.reshape(-1, 3)
    lo, hi = v.min(0) - 0.5, v.max(0) + 0.5
    o = (lo + (hi - lo) * rng.random((n, 3))).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    return o, d


def _static(g, o_dev, d_dev, n, mode, out_dev):
    from eos_amd import _lib as E
    opts = E.TraceOpts(E.QUERY[mode], E.EOS_MEM_DEVICE)
    hits = out_dev if mode != "occluded" else None
    occ = out_dev if mode == "occluded" else None
    return E.lib().eos_trace_rays_static(g.h, n, o_dev, d_dev, C.byref(opts), hits, occ)


@pytest.mark.parametrize("width", [4, 2])
def test_static_queries_equal_allocating_queries(width):
    import torch
    from eos_amd import _lib as E
    s = scenes.config_f(64, 48, instances=4, nu=32, nv=16)     # transformed triangles included
    g = E.Scene(s, leaf_max=4, width=width)
    for n, seed in ((300_000, 1), (40_000, 2)):               # the second call is smaller
        o, d = _rays(s, n, seed)
        print(f"static queries width={width} n={n}", flush=True)
        to, td = torch.from_numpy(o).cuda(), torch.from_numpy(d).cuda()
        for mode in ("nearest", "occluded", "linear"):
            if mode == "linear" and n > 50_000:
                continue                                       # every ray tests every triangle
            words = n * 12 if mode != "occluded" else (n + 3) // 4
            a = torch.full((words,), -7, dtype=torch.int32, device="cuda")
            b = torch.full((words,), -7, dtype=torch.int32, device="cuda")
            torch.cuda.synchronize()
            E.check(_static(g, to.data_ptr(), td.data_ptr(), n, mode, a.data_ptr()))
            g.trace_rays_device(n, to.data_ptr(), td.data_ptr(), mode,
                                hits_ptr=b.data_ptr() if mode != "occluded" else None,
                                occluded_ptr=b.data_ptr() if mode == "occluded" else None)
            torch.cuda.synchronize()
            assert torch.equal(a, b), mode
            if mode == "nearest":
                rec = a.cpu().numpy().view(E.RAY_HIT)
                assert (rec["triangle"] >= 0).sum() > n // 50


def test_static_query_limits():
    import torch
    from eos_amd import _lib as E
    s = scenes.config_a(16, 16)
    g = E.Scene(s, leaf_max=4)
    o = torch.zeros((8, 3), device="cuda")
    out = torch.zeros(8 * 12, dtype=torch.int32, device="cuda")
    st = _static(g, o.data_ptr(), o.data_ptr(), E.STATIC_MAX_RAYS + 1, "nearest", out.data_ptr())
    assert st == E.EOS_ERR_LIMIT and b"at most" in E.lib().eos_last_error()
    host = np.zeros((8, 3), np.float32)
    opts = E.TraceOpts(E.QUERY["nearest"], E.EOS_MEM_HOST)
    st = E.lib().eos_trace_rays_static(g.h, 8, host.ctypes.data, host.ctypes.data, C.byref(opts),
                                       out.data_ptr(), None)
    assert st == E.EOS_ERR_INVALID and b"device memory" in E.lib().eos_last_error()
    assert _static(g, o.data_ptr(), o.data_ptr(), 0, "nearest", out.data_ptr()) == E.EOS_OK
