"""ctypes binding of the C ABI in include/eos_c.h (libeos_amd.so, built in-tree).

There is no fallback: if the HIP library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libeos_amd.so")

EOS_OK, EOS_ERR_INVALID, EOS_ERR_HIP, EOS_ERR_NOMEM, EOS_ERR_STATE, EOS_ERR_LIMIT, EOS_ERR_NODEVICE = range(7)
ABI_VERSION = 7   # EOS_ABI_VERSION of include/eos_c.h these structures follow
EOS_MEM_HOST, EOS_MEM_DEVICE = 0, 1
DBG = {"morton_sorted": 1, "perm": 2, "kchild": 3, "kparent": 4, "ibox": 5, "nodes": 6, "tris": 7,
       "morton": 8, "nodes4": 9, "xforms": 10}

# Every symbol include/eos_c.h declares.
EXPORTS = ["eos_create_scene", "eos_destroy_scene", "eos_build_lbvh", "eos_build_lbvh_ex", "eos_update_materials",
           "eos_update_lights", "eos_render", "eos_render_adjoint", "eos_render_batch",
           "eos_set_launch_config", "eos_get_launch_config", "eos_debug_read", "eos_last_error",
           "eos_abi_version", "eos_device_name", "eos_mesh_load_obj", "eos_mesh_free", "eos_set_transforms",
           "eos_set_ambient_refractive_index", "eos_render_frames", "eos_render_rows", "eos_trace_rays",
           "eos_trace_rays_static", "eos_build_lbvh_batch"]
STATIC_MAX_RAYS = 1 << 20   # EOS_STATIC_MAX_RAYS

# eos_trace_rays query modes and the eos_ray_hit record (48 bytes)
QUERY = {"nearest": 0, "occluded": 1, "linear": 2}
RAY_HIT = np.dtype([("point", np.float32, 3), ("hit", np.int32), ("normal", np.float32, 3), ("material", np.int32),
                    ("t", np.float32), ("u", np.float32), ("v", np.float32), ("triangle", np.int32)])
assert RAY_HIT.itemsize == 48


class EosError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"eos status {status}: {msg}")
        self.status = status


class SceneDesc(C.Structure):
    _fields_ = [("num_triangles", C.c_int32), ("vertices", C.c_void_p), ("normals", C.c_void_p),
                ("material_index", C.c_void_p), ("triangle_memory", C.c_int32),
                ("num_materials", C.c_int32), ("materials", C.c_void_p),
                ("num_lights", C.c_int32), ("lights", C.c_void_p)]


class BuildStats(C.Structure):
    _fields_ = [("build_ms", C.c_double), ("num_internal", C.c_int32), ("root_ref", C.c_int32),
                ("leaf_max", C.c_int32), ("width", C.c_int32), ("num_nodes4", C.c_int32), ("root4", C.c_int32)]


class BuildOpts(C.Structure):
    _fields_ = [("leaf_max", C.c_int32), ("width", C.c_int32)]


class RenderOpts(C.Structure):
    _fields_ = [("depth", C.c_int32), ("output_memory", C.c_int32), ("rgb_out", C.c_void_p),
                ("hit_out", C.c_void_p), ("keep_records", C.c_int32), ("count_visits", C.c_int32),
                ("aa_samples", C.c_int32), ("launch_times", C.c_int32)]


class RenderStats(C.Structure):
    _fields_ = [("primary_rays", C.c_int64), ("shadow_rays", C.c_int64), ("secondary_rays", C.c_int64),
                ("hits", C.c_int64), ("node_visits", C.c_int64), ("tri_tests", C.c_int64),
                ("render_ms", C.c_double), ("trace_closest_ms", C.c_double),
                ("trace_shadow_ms", C.c_double), ("shade_ms", C.c_double),
                ("trace_closest_launches", C.c_int32), ("trace_shadow_launches", C.c_int32),
                ("closest_visits_first", C.c_int64), ("closest_tests_first", C.c_int64),
                ("trace_first_ms", C.c_double), ("primary_culled", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class AdjointOpts(C.Structure):
    _fields_ = [("depth", C.c_int32), ("memory", C.c_int32), ("recompute", C.c_int32), ("pad", C.c_int32)]


class AdjointStats(C.Structure):
    _fields_ = [("adjoint_ms", C.c_double), ("retrace_ms", C.c_double), ("records", C.c_int64)]


class BatchOpts(C.Structure):
    _fields_ = [("depth", C.c_int32), ("output_memory", C.c_int32), ("use_graph", C.c_int32),
                ("pad", C.c_int32)]


class TraceOpts(C.Structure):
    _fields_ = [("mode", C.c_int32), ("memory", C.c_int32)]


class BatchStats(C.Structure):
    _fields_ = [("render_ms", C.c_double), ("rays", C.c_int64)]


class Mesh(C.Structure):
    _fields_ = [("num_triangles", C.c_int32), ("vertices", C.POINTER(C.c_float)),
                ("normals", C.POINTER(C.c_float))]


class LaunchConfig(C.Structure):
    _fields_ = [("workgroup", C.c_int32), ("stack_lds", C.c_int32), ("waves_per_cu", C.c_int32),
                ("refill", C.c_int32), ("batch", C.c_int32), ("primary_packets", C.c_int32), ("coop", C.c_int32),
                ("frames_waves_per_cu", C.c_int32), ("frames_refill", C.c_int32), ("frames_batch", C.c_int32),
                ("group", C.c_int32), ("steal", C.c_int32), ("scatter", C.c_int32),
                ("group_lanes", C.c_int32), ("cull", C.c_int32)]


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python eos_amd/build.py` "
                               "(there is no CPU fallback)")
        # torch ships its own libamdhip64 (same soname).  Loaded after this
        # library, it would be a second HIP runtime in the process, and that one
        # sees no GPU; loaded first, this library binds to it and the process has
        # one runtime, so device pointers pass freely between the two.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        vp, i32 = C.c_void_p, C.c_int32
        L.eos_create_scene.argtypes = [vp, C.POINTER(vp)]
        L.eos_destroy_scene.argtypes = [vp]
        L.eos_build_lbvh.argtypes = [vp, i32, vp]
        L.eos_build_lbvh_ex.argtypes = [vp, vp, vp]
        L.eos_update_materials.argtypes = [vp, vp, i32]
        L.eos_update_lights.argtypes = [vp, vp, i32]
        L.eos_render.argtypes = [vp, vp, vp, vp, vp]
        L.eos_render_adjoint.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.eos_render_batch.argtypes = [vp, vp, i32, vp, vp, vp]
        L.eos_set_launch_config.argtypes = [vp, vp]
        L.eos_get_launch_config.argtypes = [vp, vp]
        L.eos_debug_read.argtypes = [vp, i32, vp, C.c_int64]
        L.eos_last_error.restype = C.c_char_p
        L.eos_abi_version.restype = i32
        L.eos_device_name.argtypes = [C.c_char_p, i32]
        L.eos_mesh_load_obj.argtypes = [C.c_char_p, vp]
        L.eos_mesh_free.argtypes = [vp]
        L.eos_set_transforms.argtypes = [vp, i32, vp, vp, i32]
        L.eos_set_ambient_refractive_index.argtypes = [vp, C.c_float]
        L.eos_render_frames.argtypes = [vp, vp, i32, vp, vp, vp]
        L.eos_render_rows.argtypes = [vp, vp, i32, i32, vp, vp, vp]
        L.eos_trace_rays.argtypes = [vp, i32, vp, vp, vp, vp, vp]
        L.eos_trace_rays_static.argtypes = [vp, i32, vp, vp, vp, vp, vp]
        L.eos_build_lbvh_batch.argtypes = [vp, i32, vp, vp]
        for name in EXPORTS:
            if name != "eos_last_error" and name != "eos_abi_version":
                getattr(L, name).restype = C.c_int
        if L.eos_abi_version() != ABI_VERSION:
            raise RuntimeError(f"{LIB_PATH}: ABI version {L.eos_abi_version()}, these bindings need {ABI_VERSION}; "
                               "rebuild the library")
        _LIB = L
    return _LIB


def check(status: int):
    if status != EOS_OK:
        raise EosError(status, lib().eos_last_error().decode())


def device_name() -> str:
    buf = C.create_string_buffer(256)
    check(lib().eos_device_name(buf, 256))
    return buf.value.decode()


def load_obj(path: str):
    """Triangles of a Wavefront OBJ file (eos_mesh_load_obj): vertices and
    per-corner normals, float32 [N][3][3] each."""
    m = Mesh()
    check(lib().eos_mesh_load_obj(os.fsencode(path), C.byref(m)))
    try:
        n = m.num_triangles
        v = np.ctypeslib.as_array(m.vertices, shape=(n * 9,)).reshape(n, 3, 3).copy() if n else np.zeros((0, 3, 3), np.float32)
        nn = np.ctypeslib.as_array(m.normals, shape=(n * 9,)).reshape(n, 3, 3).copy() if n else np.zeros((0, 3, 3), np.float32)
    finally:
        lib().eos_mesh_free(C.byref(m))
    return v, nn


def _ptr(a):
    return None if a is None else a.ctypes.data


class Scene:
    """One device scene (eos_scene*).  Arrays are copied to the device at creation."""

    def __init__(self, scene, leaf_max: int = 4, build: bool = True, width: int = 4):
        self.data = scene
        v = np.ascontiguousarray(scene.verts, np.float32)
        n = np.ascontiguousarray(scene.normals, np.float32)
        m = np.ascontiguousarray(scene.mat, np.int32)
        self._mats = np.ascontiguousarray(scene.materials)
        self._lights = np.ascontiguousarray(scene.lights)
        d = SceneDesc(v.shape[0], _ptr(v), _ptr(n), _ptr(m), EOS_MEM_HOST, len(self._mats),
                      _ptr(self._mats), len(self._lights), _ptr(self._lights))
        h = C.c_void_p()
        check(lib().eos_create_scene(C.byref(d), C.byref(h)))
        self.h = h
        self.build_stats = None
        # bumped by every call that changes the scene or replaces its kept hit records
        self.generation = 0
        xf = getattr(scene, "xforms", None)
        if xf is not None:
            self.set_transforms(xf, scene.xform_index)
        amb = float(getattr(scene, "ambient_index", 1.0))
        if amb != 1.0:
            check(lib().eos_set_ambient_refractive_index(self.h, amb))
        if build:
            self.build(leaf_max, width)

    @property
    def handle(self):
        return self.h

    def set_transforms(self, xforms, xform_index):
        """eos_set_transforms: [T][16] column-major float32 matrices and a per-triangle
        index (-1 = world space).  The scene must be built again before rendering."""
        self.generation += 1
        self._xf = np.ascontiguousarray(xforms, np.float32).reshape(-1, 16)
        self._xi = np.ascontiguousarray(xform_index, np.int32)
        check(lib().eos_set_transforms(self.h, len(self._xf), _ptr(self._xf), _ptr(self._xi), EOS_MEM_HOST))
        self.n_xforms = len(self._xf)

    def build(self, leaf_max: int = 4, width: int = 4) -> dict:
        self.generation += 1
        st = BuildStats()
        o = BuildOpts(leaf_max, width)
        check(lib().eos_build_lbvh_ex(self.h, C.byref(o), C.byref(st)))
        self.build_stats = {k: getattr(st, k) for k, _ in st._fields_}
        return self.build_stats

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().eos_destroy_scene(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_materials(self, mats):
        self.generation += 1
        self._mats = np.ascontiguousarray(mats)
        check(lib().eos_update_materials(self.h, _ptr(self._mats), len(self._mats)))

    def set_lights(self, lights):
        self.generation += 1
        self._lights = np.ascontiguousarray(lights)
        check(lib().eos_update_lights(self.h, _ptr(self._lights), len(self._lights)))

    DEFAULT_LAUNCH = {"workgroup": 256, "stack_lds": 16, "waves_per_cu": 20, "refill": 16, "batch": 64,
                      "primary_packets": 0, "coop": 0, "frames_waves_per_cu": 16, "frames_refill": 16,
                      "frames_batch": 64, "group": 1, "steal": 1, "scatter": 4, "group_lanes": 64,
                      "cull": 1}

    def set_launch_config(self, *args, **kw):
        """Positional fields in eos_launch_config order, or keywords; fields not
        given keep their current value.  No arguments: the defaults."""
        cfg = dict(self.DEFAULT_LAUNCH) if not args and not kw else self.launch_config()
        cfg.update(dict(zip([f for f, _ in LaunchConfig._fields_], args)))
        cfg.update(kw)
        c = LaunchConfig(**cfg)
        check(lib().eos_set_launch_config(self.h, C.byref(c)))

    def launch_config(self) -> dict:
        c = LaunchConfig()
        check(lib().eos_get_launch_config(self.h, C.byref(c)))
        return {k: getattr(c, k) for k, _ in c._fields_}

    def render(self, camera=None, depth=None, want_rgb=False, want_hits=False, keep_records=False,
               count_visits=False, want_rgba=True, aa_samples=0, launch_times=False):
        self.generation += 1
        cam = np.array(self.data.camera if camera is None else camera, copy=True)
        depth = self.data.depth if depth is None else depth
        W, H = int(cam["width"]), int(cam["height"])
        rgba = np.zeros((H, W, 4), np.uint8) if want_rgba else None
        rgb = np.zeros((H, W, 3), np.float32) if want_rgb else None
        hits = np.zeros((H, W), np.int32) if want_hits else None
        o = RenderOpts(depth, EOS_MEM_HOST, _ptr(rgb), _ptr(hits), int(keep_records), int(count_visits),
                       int(aa_samples), int(launch_times))
        st = RenderStats()
        check(lib().eos_render(self.h, cam.ctypes.data, C.byref(o), _ptr(rgba), C.byref(st)))
        return {"rgba": rgba, "rgb": rgb, "hit": hits, "stats": st.as_dict()}

    def render_device(self, camera, depth, rgba_dev_ptr, rgb_dev_ptr=None, hit_dev_ptr=None,
                      keep_records=False, count_visits=False, aa_samples=0, launch_times=False) -> dict:
        """Render into device buffers (raw pointers, e.g. torch tensors' data_ptr()).
        launch_times: also time every launch (trace_closest_ms, trace_shadow_ms, shade_ms)."""
        self.generation += 1
        cam = np.array(camera, copy=True)
        o = RenderOpts(depth, EOS_MEM_DEVICE, rgb_dev_ptr, hit_dev_ptr, int(keep_records), int(count_visits),
                       int(aa_samples), int(launch_times))
        st = RenderStats()
        check(lib().eos_render(self.h, cam.ctypes.data, C.byref(o), rgba_dev_ptr, C.byref(st)))
        return st.as_dict()

    def render_rows(self, row_begin, row_end, camera=None, depth=None, want_rgb=False, want_hits=False,
                    aa_samples=0):
        """eos_render_rows: rows [row_begin, row_end) of the frame (host outputs for the band)."""
        self.generation += 1
        cam = np.array(self.data.camera if camera is None else camera, copy=True)
        depth = self.data.depth if depth is None else depth
        W, Hb = int(cam["width"]), max(row_end - row_begin, 0)
        rgba = np.zeros((Hb, W, 4), np.uint8)
        rgb = np.zeros((Hb, W, 3), np.float32) if want_rgb else None
        hits = np.zeros((Hb, W), np.int32) if want_hits else None
        o = RenderOpts(depth, EOS_MEM_HOST, _ptr(rgb), _ptr(hits), 0, 0, int(aa_samples))
        st = RenderStats()
        check(lib().eos_render_rows(self.h, cam.ctypes.data, row_begin, row_end, C.byref(o), _ptr(rgba), C.byref(st)))
        return {"rgba": rgba, "rgb": rgb, "hit": hits, "stats": st.as_dict()}

    def render_frames(self, cameras, rgba_dev_ptrs, depth=None, count_visits=False, launch_times=False) -> list:
        """eos_render_frames: frame i from cameras[i] into the device buffer rgba_dev_ptrs[i]
        (two frames in flight); per-frame stats (launch_times: with each primary-ray launch timed)."""
        self.generation += 1
        cams = np.ascontiguousarray(cameras)
        n = len(cams)
        depth = self.data.depth if depth is None else depth
        o = RenderOpts(depth, EOS_MEM_DEVICE, None, None, 0, int(count_visits), 0, int(launch_times))
        ptrs = (C.c_void_p * n)(*rgba_dev_ptrs)
        st = (RenderStats * n)()
        check(lib().eos_render_frames(self.h, cams.ctypes.data, n, C.byref(o), ptrs, st))
        return [s.as_dict() for s in st]

    def trace_rays(self, origins, directions, mode: str = "nearest"):
        """eos_trace_rays on host arrays [n][3]: for "nearest" (Scene::findMinDistanceIntersectionBVH)
        and "linear" (Scene::findMinDistanceIntersectionLinear) a RAY_HIT record per ray, for
        "occluded" (Scene::shadowRayVisibilityBVH) a bool per ray."""
        o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(directions, np.float32).reshape(-1, 3)
        if len(o) != len(d):
            raise ValueError("origins and directions differ in length")
        n = len(o)
        opts = TraceOpts(QUERY[mode], EOS_MEM_HOST)
        if mode == "occluded":
            occ = np.zeros(n, np.uint8)
            check(lib().eos_trace_rays(self.h, n, _ptr(o), _ptr(d), C.byref(opts), None, _ptr(occ)))
            return occ.astype(bool)
        hits = np.zeros(n, RAY_HIT)
        check(lib().eos_trace_rays(self.h, n, _ptr(o), _ptr(d), C.byref(opts), _ptr(hits), None))
        return hits

    def trace_rays_device(self, n, origins_ptr, directions_ptr, mode: str = "nearest", hits_ptr=None,
                          occluded_ptr=None):
        """eos_trace_rays with device buffers (raw pointers): origins / directions [n][3] float32,
        hits [n] RAY_HIT records or occluded [n] uint8."""
        opts = TraceOpts(QUERY[mode], EOS_MEM_DEVICE)
        check(lib().eos_trace_rays(self.h, n, origins_ptr, directions_ptr, C.byref(opts), hits_ptr, occluded_ptr))

    def adjoint(self, dl_drgb, camera=None, depth=None, recompute=False):
        cam = np.array(self.data.camera if camera is None else camera, copy=True)
        depth = self.data.depth if depth is None else depth
        g = np.ascontiguousarray(dl_drgb, np.float32)
        gm = np.zeros((len(self._mats), 7), np.float64)
        gl = np.zeros((max(len(self._lights), 1), 6), np.float64)
        o = AdjointOpts(depth, EOS_MEM_HOST, int(recompute), 0)
        st = AdjointStats()
        check(lib().eos_render_adjoint(self.h, cam.ctypes.data, C.byref(o), _ptr(g), _ptr(gm), _ptr(gl),
                                       C.byref(st)))
        return gm, gl[:len(self._lights)], {k: getattr(st, k) for k, _ in st._fields_}

    def adjoint_device(self, camera, depth, dl_dev_ptr, gm_dev_ptr, gl_dev_ptr, recompute=False) -> dict:
        """Adjoint with device buffers: dl [H][W][3] float32, gm [M][7] and gl [L][6] float64 (accumulated)."""
        cam = np.array(camera, copy=True)
        o = AdjointOpts(depth, EOS_MEM_DEVICE, int(recompute), 0)
        st = AdjointStats()
        check(lib().eos_render_adjoint(self.h, cam.ctypes.data, C.byref(o), dl_dev_ptr, gm_dev_ptr, gl_dev_ptr,
                                       C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def debug_read(self, what: str):
        n = self.data.n_tri
        m = max(n - 1, 0)
        shapes = {"morton_sorted": ((n,), np.uint32), "perm": ((n,), np.int32),
                  "kchild": ((m, 2), np.int32), "kparent": ((max(2 * n - 1, 0),), np.int32),
                  "ibox": ((m, 6), np.float32), "nodes": ((m, 16), np.float32),
                  "tris": ((n, 12), np.float32), "morton": ((n,), np.uint32),
                  "nodes4": ((self.build_stats["num_nodes4"] if self.build_stats else 0, 32), np.float32),
                  "xforms": ((getattr(self, "n_xforms", 0), 36), np.float32)}
        shape, dt = shapes[what]
        out = np.zeros(shape, dt)
        check(lib().eos_debug_read(self.h, DBG[what], _ptr(out), out.nbytes))
        return out


def build_batch(scenes, leaf_max: int = 4, width: int = 4) -> list:
    """eos_build_lbvh_batch: the LBVHs of every scene in one launch sequence; per-scene build stats."""
    hs = (C.c_void_p * len(scenes))(*[s.h.value for s in scenes])
    st = (BuildStats * len(scenes))()
    o = BuildOpts(leaf_max, width)
    check(lib().eos_build_lbvh_batch(hs, len(scenes), C.byref(o), st))
    out = []
    for s, b in zip(scenes, st):
        s.generation += 1
        s.build_stats = {k: getattr(b, k) for k, _ in b._fields_}
        out.append(s.build_stats)
    return out


def render_batch(scenes, cameras=None, depth=2, use_graph=False, device_ptrs=None):
    """Renders all scenes in one batched frame.  Outputs go to host arrays (returned),
    or to the device buffers in ``device_ptrs`` (one RGBA8 image per scene)."""
    cams = np.ascontiguousarray(np.stack([s.data.camera for s in scenes]) if cameras is None else cameras)
    hs = (C.c_void_p * len(scenes))(*[s.h.value for s in scenes])
    if device_ptrs is None:
        outs = [np.zeros((int(c["height"]), int(c["width"]), 4), np.uint8) for c in cams]
        ptrs = (C.c_void_p * len(scenes))(*[o.ctypes.data for o in outs])
        mem = EOS_MEM_HOST
    else:
        outs = None
        ptrs = (C.c_void_p * len(scenes))(*device_ptrs)
        mem = EOS_MEM_DEVICE
    o = BatchOpts(depth, mem, int(use_graph), 0)
    st = BatchStats()
    check(lib().eos_render_batch(hs, cams.ctypes.data, len(scenes), C.byref(o), ptrs, C.byref(st)))
    return outs, {k: getattr(st, k) for k, _ in st._fields_}
