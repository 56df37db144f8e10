"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product library never loads it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

REC_DTYPE = np.dtype([("tri", "<i4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4"),
                      ("p", "<f4", (3,)), ("n", "<f4", (3,)), ("mat", "<i4"),
                      ("vis", "<u4"), ("refl", "<f4"), ("pad", "<f4", (3,))])
assert REC_DTYPE.itemsize == 64


class OrcScene(C.Structure):
    _fields_ = [("n_tri", C.c_int32), ("verts", C.c_void_p), ("normals", C.c_void_p),
                ("mat", C.c_void_p), ("n_mat", C.c_int32), ("mats", C.c_void_p),
                ("n_light", C.c_int32), ("lights", C.c_void_p),
                ("n_xform", C.c_int32), ("xforms", C.c_void_p), ("xform_index", C.c_void_p),
                ("ambient_index", C.c_float)]


def _xform_arrays(scene):
    """(n_xform, xforms [T][16] float32, xform_index [N] int32) or (0, None, None)."""
    xf = getattr(scene, "xforms", None)
    xi = getattr(scene, "xform_index", None)
    if xf is None or xi is None:
        return 0, None, None
    xf = np.ascontiguousarray(xf, np.float32).reshape(-1, 16)
    xi = np.ascontiguousarray(xi, np.int32)
    return len(xf), xf, xi


def _orc_scene(holder, scene):
    """Fills holder's arrays from the scene and returns an OrcScene over them."""
    holder.verts = np.ascontiguousarray(scene.verts, np.float32)
    holder.normals = np.ascontiguousarray(scene.normals, np.float32)
    holder.mat = np.ascontiguousarray(scene.mat, np.int32)
    holder.materials = np.ascontiguousarray(scene.materials)
    holder.lights = np.ascontiguousarray(scene.lights)
    holder.n_xform, holder.xforms, holder.xform_index = _xform_arrays(scene)
    return OrcScene(holder.verts.shape[0], _p(holder.verts), _p(holder.normals), _p(holder.mat),
                    len(holder.materials), _p(holder.materials), len(holder.lights), _p(holder.lights),
                    holder.n_xform, _p(holder.xforms), _p(holder.xform_index),
                    float(getattr(scene, "ambient_index", 1.0)))


class OrcCounts(C.Structure):
    _fields_ = [("primary", C.c_int64), ("shadow", C.c_int64), ("secondary", C.c_int64),
                ("hits", C.c_int64), ("node_visits", C.c_int64), ("tri_tests", C.c_int64)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


def build() -> None:
    subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])


_FAST = None


def lib_fast():
    """The same restatement built with the reference's timing flags (-O3
    -ffast-math, oracle/Makefile: liboracle_fast.so): for CPU timings only,
    never for parity."""
    global _FAST
    if _FAST is None:
        path = os.path.join(_HERE, "liboracle_fast.so")
        if not os.path.exists(path):
            subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle_fast.so"])
        _FAST = _bind(C.CDLL(path))
    return _FAST


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = _bind(C.CDLL(path))
    return _LIB


def _bind(L):
    vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
    L.orc_accel_build.restype = vp
    L.orc_accel_build.argtypes = [vp, C.c_int]
    L.orc_accel_free.argtypes = [vp]
    L.orc_nearest.argtypes = [vp, vp, i32, vp, vp, i32, vp, vp, vp, vp, vp, vp]
    L.orc_render.argtypes = [vp, vp, vp, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, vp]
    L.orc_nearest_linear.argtypes = [vp, i32, vp, vp, vp, vp, vp, vp, vp]
    L.orc_render_lbvh.argtypes = [vp, vp, i32, vp, vp, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, i32]
    L.orc_mt.argtypes = [vp, vp, vp, vp]
    L.orc_render_aa.argtypes = [vp, vp, vp, C.c_int, C.c_int, vp, vp, vp]
    L.orc_mat4_inverse.argtypes = [vp, vp]
    L.orc_camera_ray.argtypes = [vp, C.c_int, C.c_int, vp]
    L.orc_camera_ray_aa.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, vp]
    L.orc_quantize.restype = C.c_uint8
    L.orc_quantize.argtypes = [C.c_float]
    L.orc_morton3.restype = C.c_uint32
    L.orc_morton3.argtypes = [C.c_uint32] * 3
    L.orc_shade_phong.restype = C.c_float
    L.orc_shade_phong.argtypes = [vp, vp, C.c_int, C.c_int, vp, vp, vp]
    L.orc_morton_codes.argtypes = [i32, vp, vp, vp]
    L.orc_lbvh_build.argtypes = [i32, vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.orc_karras.argtypes = [i32, vp, vp, vp, vp, vp]
    L.orc_lbvh_trace.argtypes = [vp, i32, vp, i32, vp, vp, i32, vp, vp, vp, vp, vp, vp, i32, i32, vp]
    L.orc_lbvh4_packet_primary.argtypes = [vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, i32, vp]
    L.orc_loss64.restype = C.c_double
    L.orc_loss64.argtypes = [vp, i32, i64, vp, i32, i32, vp, vp, vp]
    L.orc_grad_fd64.argtypes = [vp, i32, i64, vp, i32, i32, vp, vp, C.c_double, vp, i32, vp]
    L.orc_grad_fwd64.argtypes = [vp, i32, i64, vp, i32, i32, vp, vp, vp]
    return L


def _p(a):
    return None if a is None else a.ctypes.data


class Oracle:
    """Holds one scene (arrays kept alive) and its exact nearest-hit accelerator."""

    def __init__(self, scene, brute_force: bool = False):
        self.scene = scene
        self._s = _orc_scene(self, scene)
        self._accel = lib().orc_accel_build(C.byref(self._s), 1 if brute_force else 0)

    def set_params(self, materials=None, lights=None):
        if materials is not None:
            self.materials = np.ascontiguousarray(materials)
            self._s.mats = _p(self.materials)
            self._s.n_mat = len(self.materials)
        if lights is not None:
            self.lights = np.ascontiguousarray(lights)
            self._s.lights = _p(self.lights)
            self._s.n_light = len(self.lights)

    def __del__(self):
        try:
            if self._accel:
                lib().orc_accel_free(self._accel)
        except Exception:
            pass

    def nearest(self, origins, dirs, any_hit=False):
        o = np.ascontiguousarray(origins, np.float32)
        d = np.ascontiguousarray(dirs, np.float32)
        n = o.shape[0]
        tri = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        u = np.zeros(n, np.float32)
        v = np.zeros(n, np.float32)
        P = np.zeros((n, 3), np.float32)
        N = np.zeros((n, 3), np.float32)
        lib().orc_nearest(C.byref(self._s), self._accel, n, _p(o), _p(d), 1 if any_hit else 0, _p(tri),
                          _p(t), _p(u), _p(v), _p(P), _p(N))
        if any_hit:
            return tri.astype(bool)
        return {"tri": tri, "t": t, "u": u, "v": v, "point": P, "normal": N}

    def nearest_linear(self, origins, dirs):
        """Scene::findMinDistanceIntersectionLinear: found flag, kept triangle, t, point, normal."""
        o = np.ascontiguousarray(origins, np.float32)
        d = np.ascontiguousarray(dirs, np.float32)
        n = o.shape[0]
        found = np.zeros(n, np.int32)
        tri = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        P = np.zeros((n, 3), np.float32)
        N = np.zeros((n, 3), np.float32)
        lib().orc_nearest_linear(C.byref(self._s), n, _p(o), _p(d), _p(found), _p(tri), _p(t), _p(P), _p(N))
        return {"found": found.astype(bool), "tri": tri, "t": t, "point": P, "normal": N}

    def render_aa(self, n, camera=None, depth=None):
        """renderWithAA: n x n samples per pixel."""
        cam = np.array(self.scene.camera if camera is None else camera, copy=True)
        depth = self.scene.depth if depth is None else depth
        W, H = int(cam["width"]), int(cam["height"])
        rgba = np.zeros((H, W, 4), np.uint8)
        rgb = np.zeros((H, W, 3), np.float32)
        cnt = OrcCounts()
        lib().orc_render_aa(C.byref(self._s), self._accel, cam.ctypes.data, depth, n, _p(rgba), _p(rgb),
                            C.byref(cnt))
        return {"rgba": rgba, "rgb": rgb, "counts": cnt.as_dict()}

    def render(self, camera=None, depth=None, row_start=0, row_step=1, want_recs=False):
        cam = np.array(self.scene.camera if camera is None else camera, copy=True)
        depth = self.scene.depth if depth is None else depth
        W, H = int(cam["width"]), int(cam["height"])
        rgba = np.zeros((H, W, 4), np.uint8)
        rgb = np.zeros((H, W, 3), np.float32)
        hits = np.full((H, W), -1, np.int32)
        recs = np.zeros((depth + 1, H, W), REC_DTYPE) if want_recs else None
        cnt = OrcCounts()
        lib().orc_render(C.byref(self._s), self._accel, cam.ctypes.data, depth, row_start, row_step,
                         _p(rgba), _p(rgb), _p(hits), _p(recs), C.byref(cnt))
        out = {"rgba": rgba, "rgb": rgb, "hit": hits, "counts": cnt.as_dict()}
        if want_recs:
            out["recs"] = recs
        return out


def morton_codes(verts):
    v = np.ascontiguousarray(verts, np.float32)
    n = v.shape[0]
    codes = np.zeros(n, np.uint32)
    cb = np.zeros(6, np.float32)
    lib().orc_morton_codes(n, _p(v), _p(codes), _p(cb))
    return codes, cb


def morton3(x, y, z) -> int:
    return int(lib().orc_morton3(x, y, z))


def karras(sorted_codes):
    k = np.ascontiguousarray(sorted_codes, np.uint32)
    n = len(k)
    kc = np.zeros(max(2 * (n - 1), 1), np.int32)
    kp = np.zeros(max(2 * n - 1, 1), np.int32)
    fi = np.zeros(max(n - 1, 1), np.int32)
    la = np.zeros(max(n - 1, 1), np.int32)
    lib().orc_karras(n, _p(k), _p(kc), _p(kp), _p(fi), _p(la))
    return kc[:2 * (n - 1)].reshape(-1, 2), kp[:2 * n - 1], fi[:n - 1], la[:n - 1]


def lbvh_build(verts, leaf_max=1, xforms=None, xform_index=None):
    """CPU LBVH restatement.  With transformations, keys and boxes come from the
    world-space vertices and the records keep the local ones."""
    v = np.ascontiguousarray(verts, np.float32)
    xf = None if xforms is None else np.ascontiguousarray(xforms, np.float32).reshape(-1, 16)
    xi = None if xform_index is None else np.ascontiguousarray(xform_index, np.int32)
    n = v.shape[0]
    m = max(n - 1, 1)
    out = {
        "codes": np.zeros(n, np.uint32), "perm": np.zeros(n, np.int32),
        "kchild": np.zeros(2 * m, np.int32), "kparent": np.zeros(2 * n - 1 if n else 1, np.int32),
        "ibox": np.zeros(6 * m, np.float32), "nodes": np.zeros(16 * m, np.float32),
        "tris": np.zeros(12 * max(n, 1), np.float32), "root": np.zeros(1, np.int32)}
    n4 = np.zeros(1, np.int32)
    r4 = np.zeros(1, np.int32)
    lib().orc_lbvh_build(n, _p(v), leaf_max, None, None, None, None, None, None, None, None, None,
                         _p(n4), None, _p(xf), _p(xi))
    out["nodes4"] = np.zeros(32 * max(int(n4[0]), 1), np.float32)
    lib().orc_lbvh_build(n, _p(v), leaf_max, _p(out["codes"]), _p(out["perm"]), _p(out["kchild"]),
                         _p(out["kparent"]), _p(out["ibox"]), _p(out["nodes"]), _p(out["tris"]),
                         _p(out["root"]), _p(out["nodes4"]), _p(n4), _p(r4), _p(xf), _p(xi))
    out["nodes4"] = out["nodes4"][:32 * int(n4[0])].reshape(-1, 32)
    out["root4"] = int(r4[0])
    out["kchild"] = out["kchild"][:2 * (n - 1)].reshape(-1, 2)
    out["ibox"] = out["ibox"][:6 * (n - 1)].reshape(-1, 6)
    out["nodes"] = out["nodes"][:16 * (n - 1)].reshape(-1, 16)
    out["tris"] = out["tris"][:12 * n].reshape(-1, 12)
    out["root"] = int(out["root"][0])
    return out


def lbvh_trace(b, origins, dirs, any_hit=False, width=2, xforms=None):
    o = np.ascontiguousarray(origins, np.float32)
    d = np.ascontiguousarray(dirs, np.float32)
    n = o.shape[0]
    hit = np.zeros(n, np.int32)
    t = np.zeros(n, np.float32)
    u = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    vis = np.zeros(1, np.int64)
    tst = np.zeros(1, np.int64)
    nodes = np.ascontiguousarray(b["nodes4"] if width == 4 else b["nodes"], np.float32)
    tris = np.ascontiguousarray(b["tris"], np.float32)
    root = b["root4"] if width == 4 else b["root"]
    xf = None if xforms is None else np.ascontiguousarray(xforms, np.float32).reshape(-1, 16)
    lib().orc_lbvh_trace(_p(nodes), root, _p(tris), n, _p(o), _p(d), 1 if any_hit else 0,
                         _p(hit), _p(t), _p(u), _p(v), _p(vis), _p(tst), width, 0 if xf is None else len(xf),
                         _p(xf))
    return {"hit": hit, "t": t, "u": u, "v": v, "visits": int(vis[0]), "tests": int(tst[0])}


def packet_primary(scene, b, camera=None):
    """Primary rays of a frame walked as 8x8 packets over the 4-wide tree (the
    GPU packet kernel's order): per-pixel hits and per-tile visit/test totals."""
    cam = np.array(scene.camera if camera is None else camera, copy=True)
    W, H = int(cam["width"]), int(cam["height"])
    hit = np.full((H, W), -1, np.int32)
    t = np.zeros((H, W), np.float32)
    u = np.zeros((H, W), np.float32)
    v = np.zeros((H, W), np.float32)
    vis = np.zeros(1, np.int64)
    tst = np.zeros(1, np.int64)
    nodes = np.ascontiguousarray(b["nodes4"], np.float32)
    tris = np.ascontiguousarray(b["tris"], np.float32)
    nx, xf, _ = _xform_arrays(scene)
    lib().orc_lbvh4_packet_primary(_p(nodes), b["root4"], _p(tris), cam.ctypes.data, _p(hit), _p(t), _p(u), _p(v),
                                   _p(vis), _p(tst), nx, _p(xf))
    return {"hit": hit, "t": t, "u": u, "v": v, "visits": int(vis[0]), "tests": int(tst[0])}


def mat4_inverse(m):
    """glm 0.9.6 inverse of a column-major 4x4 float32 matrix ([16] or [4][4]
    with m[c][r]), in glm's operation order."""
    a = np.ascontiguousarray(m, np.float32).reshape(16)
    out = np.zeros(16, np.float32)
    lib().orc_mat4_inverse(_p(a), _p(out))
    return out.reshape(np.shape(m))


def camera_rays(scene, camera=None):
    """Camera ray directions [H][W][3] (the reference formula)."""
    cam = np.array(scene.camera if camera is None else camera, copy=True)
    W, H = int(cam["width"]), int(cam["height"])
    out = np.zeros((H, W, 3), np.float32)
    for y in range(H):
        for x in range(W):
            lib().orc_camera_ray(cam.ctypes.data, x, y, out[y, x].ctypes.data)
    return out


def render_lbvh(scene, b, camera=None, depth=None, row_start=0, row_step=1, width=2, fast=False):
    """The restatement of the GPU's LBVH walk and shading; fast: the timing build (lib_fast)."""
    o = Oracle.__new__(Oracle)
    o._accel = None
    s = _orc_scene(o, scene)
    cam = np.array(scene.camera if camera is None else camera, copy=True)
    depth = scene.depth if depth is None else depth
    W, H = int(cam["width"]), int(cam["height"])
    rgba = np.zeros((H, W, 4), np.uint8)
    rgb = np.zeros((H, W, 3), np.float32)
    hits = np.full((H, W), -1, np.int32)
    cnt = OrcCounts()
    nodes = np.ascontiguousarray(b["nodes4"] if width == 4 else b["nodes"], np.float32)
    tris = np.ascontiguousarray(b["tris"], np.float32)
    root = b["root4"] if width == 4 else b["root"]
    (lib_fast() if fast else lib()).orc_render_lbvh(C.byref(s), _p(nodes), root, _p(tris), cam.ctypes.data, depth,
                          row_start, row_step, _p(rgba), _p(rgb), _p(hits), C.byref(cnt), width)
    del o
    return {"rgba": rgba, "rgb": rgb, "hit": hits, "counts": cnt.as_dict()}


def params_vector(materials, lights) -> np.ndarray:
    """Flatten materials/lights into the gradient parameter layout (float64)."""
    m = len(materials)
    nl = len(lights)
    p = np.zeros(7 * m + 6 * nl)
    pm = p[:7 * m].reshape(m, 7)
    pm[:, 0:3] = materials["kd"][:, :3]
    pm[:, 3:6] = materials["ks"][:, :3]
    pm[:, 6] = materials["shininess"]
    pl = p[7 * m:].reshape(nl, 6)
    pl[:, 0:3] = lights["position"][:, :3]
    pl[:, 3:6] = lights["color"][:, :3]
    return p


def loss64(recs, cam_pos, n_mat, n_light, params, target_rgb):
    r = np.ascontiguousarray(recs)
    lev = r.shape[0]
    npix = r[0].size
    cp = np.ascontiguousarray(cam_pos, np.float32)
    pr = np.ascontiguousarray(params, np.float64)
    tg = np.ascontiguousarray(target_rgb, np.float32)
    img = np.zeros((npix, 3))
    loss = lib().orc_loss64(_p(r), lev, npix, _p(cp), n_mat, n_light, _p(pr), _p(tg), _p(img))
    return loss, img


def grad_fd64(recs, cam_pos, n_mat, n_light, params, target_rgb, which, h=1e-3):
    r = np.ascontiguousarray(recs)
    w = np.ascontiguousarray(which, np.int32)
    out = np.zeros(len(w))
    cp = np.ascontiguousarray(cam_pos, np.float32)
    pr = np.ascontiguousarray(params, np.float64)
    tg = np.ascontiguousarray(target_rgb, np.float32)
    lib().orc_grad_fd64(_p(r), r.shape[0], r[0].size, _p(cp), n_mat, n_light, _p(pr), _p(tg), h,
                        _p(w), len(w), _p(out))
    return out


def grad_fwd64(recs, cam_pos, n_mat, n_light, params, target_rgb):
    r = np.ascontiguousarray(recs)
    cp = np.ascontiguousarray(cam_pos, np.float32)
    pr = np.ascontiguousarray(params, np.float64)
    tg = np.ascontiguousarray(target_rgb, np.float32)
    out = np.zeros(len(pr))
    lib().orc_grad_fwd64(_p(r), r.shape[0], r[0].size, _p(cp), n_mat, n_light, _p(pr), _p(tg), _p(out))
    return out


# --------------------------------------------------------------------- refport
class RpCounts(C.Structure):
    _fields_ = [("primary", C.c_int64), ("shadow", C.c_int64), ("secondary", C.c_int64),
                ("hits", C.c_int64), ("seconds", C.c_double), ("threads", C.c_int32), ("pad", C.c_int32)]


_RP = {}


def refport_lib(parity: bool = False):
    name = "librefport_parity.so" if parity else "librefport.so"
    if name not in _RP:
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            subprocess.check_call(["make", "-s", "-C", _HERE, name])
        L = C.CDLL(path)
        vp = C.c_void_p
        L.rp_build.restype = vp
        L.rp_build.argtypes = [vp, vp]
        L.rp_free.argtypes = [vp]
        L.rp_num_nodes.argtypes = [vp]
        L.rp_num_nodes.restype = C.c_int32
        L.rp_render.argtypes = [vp, vp, C.c_int, vp, C.c_int64, C.c_int32, vp, C.c_int32, vp]
        _RP[name] = L
    return _RP[name]


class RefPort:
    """Timing-faithful port of the reference CPU path (oracle/refport.h)."""

    def __init__(self, scene, parity: bool = False):
        self.L = refport_lib(parity)
        self.scene = scene
        self._s = _orc_scene(self, scene)
        bs = np.zeros(1, np.float64)
        self.h = self.L.rp_build(C.byref(self._s), bs.ctypes.data)
        self.build_seconds = float(bs[0])

    def __del__(self):
        try:
            self.L.rp_free(self.h)
        except Exception:
            pass

    def render(self, pixels=None, threads=0, camera=None, depth=None):
        cam = np.array(self.scene.camera if camera is None else camera, copy=True)
        depth = self.scene.depth if depth is None else depth
        W, H = int(cam["width"]), int(cam["height"])
        px = np.arange(W * H, dtype=np.int32) if pixels is None else np.ascontiguousarray(pixels, np.int32)
        rgba = np.zeros((len(px), 4), np.uint8)
        cnt = RpCounts()
        self.L.rp_render(self.h, cam.ctypes.data, depth, px.ctypes.data, len(px), W, rgba.ctypes.data,
                         threads, C.byref(cnt))
        counts = {k: getattr(cnt, k) for k, _ in cnt._fields_}
        return rgba, counts
