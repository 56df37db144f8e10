"""Seeded procedural scenes for the five rendering workloads.

Every scene is produced from SplitMix64 streams (a counter-based generator, so
numpy evaluates it vectorised and the output does not depend on the platform's
``random`` implementation).  Uniform floats take the top 24 bits of a draw;
normal deviates use double-precision Box-Muller.

A scene is plain arrays in the layout the C ABI (``include/eos_c.h``) takes:

* ``verts``   float32 [N, 3, 3]  triangle vertices v1, v2, v3
  (reference ``Triangle`` m_V1..m_V3, src/Triangle.h:73-75)
* ``normals`` float32 [N, 3, 3]  per-vertex normals n1, n2, n3 (m_N1..m_N3)
* ``mat``     int32   [N]        material index per triangle (m_MaterialIndex)
* ``materials`` MATERIAL_DTYPE [M]  kd, ks, shininess, reflectivity, reflective
  (reference ``Material``, src/Material.h:43-78)
* ``lights``  LIGHT_DTYPE [L]    position (w = 1) and colour
  (reference ``LightSource``, src/LightSource.h:27-39)
* ``camera``  CAMERA_DTYPE       position, viewing direction, right, up, W, H
  (reference ``Camera``, src/Camera.h:55-63; no field of view: the lengths of
  ``right`` and ``up`` set it, as in src/RayTracer.cpp:135-141)

* ``xforms``  float32 [T, 16]   local-to-world matrices, column-major (glm
  layout); ``xform_index`` int32 [N], -1 = world space (reference
  ``Scene::addTransformation`` + ``Triangle::setTransformationIndex``,
  src/Scene.cpp:96-105, src/Triangle.h:59).  None in configs A-E; config F
  instances one mesh many times through them.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

MATERIAL_DTYPE = np.dtype([
    ("kd", "<f4", (4,)), ("ks", "<f4", (4,)), ("shininess", "<f4"),
    ("reflectivity", "<f4"), ("reflective", "<i4"), ("transparent", "<i4"), ("refractive_index", "<f4"),
    ("pad", "<i4", (3,))])
LIGHT_DTYPE = np.dtype([("position", "<f4", (4,)), ("color", "<f4", (4,))])
CAMERA_DTYPE = np.dtype([
    ("position", "<f4", (3,)), ("view", "<f4", (3,)), ("right", "<f4", (3,)),
    ("up", "<f4", (3,)), ("width", "<i4"), ("height", "<i4")])
assert MATERIAL_DTYPE.itemsize == 64 and LIGHT_DTYPE.itemsize == 32
assert CAMERA_DTYPE.itemsize == 56

BASE_SEED = 0x456F73          # "Eos"
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """Draws ``offset+1 .. offset+n`` of the SplitMix64 sequence started at ``seed``."""
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def mix64(z: np.ndarray) -> np.ndarray:
    """The SplitMix64 finaliser applied to arbitrary 64-bit keys."""
    with np.errstate(over="ignore"):
        z = z.astype(np.uint64) + _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


class Rng:
    """A SplitMix64 stream; each call consumes draws in order."""

    def __init__(self, seed: int):
        self.seed = seed & 0xFFFFFFFFFFFFFFFF
        self.pos = 0

    def bits(self, n: int) -> np.ndarray:
        out = splitmix64(self.seed, n, self.pos)
        self.pos += n
        return out

    def uniform(self, n: int) -> np.ndarray:
        """U[0,1) from the top 24 bits (exact in float32)."""
        return (self.bits(n) >> np.uint64(40)).astype(np.float64) * (1.0 / (1 << 24))

    def normal(self, n: int) -> np.ndarray:
        """N(0,1) by Box-Muller in float64; each pair of draws yields cos and sin branches."""
        m = (n + 1) // 2
        b = self.bits(2 * m)
        u1 = ((b[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / (1 << 53))
        u2 = (b[1::2] >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
        r = np.sqrt(-2.0 * np.log(u1))
        out = np.empty(2 * m)
        out[0::2] = r * np.cos(2.0 * math.pi * u2)
        out[1::2] = r * np.sin(2.0 * math.pi * u2)
        return out[:n]


@dataclass
class SceneData:
    verts: np.ndarray
    normals: np.ndarray
    mat: np.ndarray
    materials: np.ndarray
    lights: np.ndarray
    camera: np.ndarray
    depth: int
    name: str = ""
    info: dict = field(default_factory=dict)
    # per-triangle local-to-world transformations: [T][16] float32 column-major
    # (glm layout) and [N] int32 indices, -1 = none; None: no transformations
    xforms: np.ndarray = None
    xform_index: np.ndarray = None
    # Scene::setAmbientRefractiveIndex: the medium outside transparent materials
    ambient_index: float = 1.0

    @property
    def n_tri(self) -> int:
        return int(self.verts.shape[0])

    @property
    def width(self) -> int:
        return int(self.camera["width"])

    @property
    def height(self) -> int:
        return int(self.camera["height"])

    def with_camera(self, width: int, height: int) -> "SceneData":
        """Same scene, camera re-aimed for a new resolution (same vertical FOV)."""
        cam = self.camera.copy()
        up_len = float(np.linalg.norm(cam["up"].astype(np.float64)))
        r = cam["right"].astype(np.float64)
        r = r / np.linalg.norm(r) * up_len * width / height
        cam["right"] = r.astype(np.float32)
        cam["width"] = width
        cam["height"] = height
        return SceneData(self.verts, self.normals, self.mat, self.materials, self.lights,
                         cam, self.depth, self.name, dict(self.info), self.xforms, self.xform_index,
                         self.ambient_index)


# ----------------------------------------------------------------- builders

def make_camera(eye, target, vfov_deg: float, width: int, height: int,
                world_up=(0.0, 1.0, 0.0)) -> np.ndarray:
    """A reference-style camera: unit view, |up| = tan(vfov/2), |right| = aspect*|up|."""
    eye = np.asarray(eye, np.float64)
    view = np.asarray(target, np.float64) - eye
    view /= np.linalg.norm(view)
    wu = np.asarray(world_up, np.float64)
    up = wu - np.dot(wu, view) * view
    up /= np.linalg.norm(up)
    right = np.cross(view, up)
    right /= np.linalg.norm(right)
    th = math.tan(math.radians(vfov_deg) * 0.5)
    cam = np.zeros((), CAMERA_DTYPE)
    cam["position"] = eye.astype(np.float32)
    cam["view"] = view.astype(np.float32)
    cam["up"] = (up * th).astype(np.float32)
    cam["right"] = (right * th * width / height).astype(np.float32)
    cam["width"] = width
    cam["height"] = height
    return cam


def make_material(kd, ks, shininess, reflectivity=0.0, reflective=False, transparent=False,
                  refractive_index=1.0) -> np.ndarray:
    m = np.zeros((), MATERIAL_DTYPE)
    m["kd"] = list(kd) + [1.0] if len(kd) == 3 else kd
    m["ks"] = list(ks) + [1.0] if len(ks) == 3 else ks
    m["shininess"] = shininess
    m["reflectivity"] = reflectivity
    m["reflective"] = 1 if reflective else 0
    m["transparent"] = 1 if transparent else 0
    m["refractive_index"] = refractive_index
    return m


def make_lights(positions, colors=None) -> np.ndarray:
    positions = np.asarray(positions, np.float64).reshape(-1, 3)
    out = np.zeros(len(positions), LIGHT_DTYPE)
    out["position"][:, :3] = positions
    out["position"][:, 3] = 1.0
    if colors is None:
        out["color"][:] = 1.0
    else:
        c = np.asarray(colors, np.float64).reshape(-1, 3)
        out["color"][:, :3] = c
        out["color"][:, 3] = 1.0
    return out


def face_normals(verts: np.ndarray) -> np.ndarray:
    v = verts.astype(np.float64)
    n = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    ln = np.linalg.norm(n, axis=1, keepdims=True)
    bad = ln[:, 0] == 0.0
    ln[bad] = 1.0
    n = n / ln
    n[bad] = (0.0, 0.0, 1.0)
    return np.repeat(n[:, None, :], 3, axis=1).astype(np.float32)


def icosphere(level: int, radius: float = 1.0, center=(0.0, 0.0, 0.0)):
    """Subdivided icosahedron: 20*4**level triangles, smooth normals n = v/|v|."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    vs = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
          (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, np.float64) / np.linalg.norm(v) for v in vs]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9),
             (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2),
             (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
             (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]
        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.array(verts)
    F = np.array(faces)
    tri = V[F]                              # [n,3,3] unit positions
    nrm = tri.copy()
    pos = tri * radius + np.asarray(center, np.float64)
    return pos.astype(np.float32), nrm.astype(np.float32)


def quad_grid(cells: int, half: float, y: float):
    """cells x cells quads on the plane y over x,z in [-half, half]; normal +y."""
    g = np.linspace(-half, half, cells + 1)
    x0, z0 = np.meshgrid(g[:-1], g[:-1], indexing="ij")
    x1, z1 = np.meshgrid(g[1:], g[1:], indexing="ij")
    x0, z0, x1, z1 = (a.ravel() for a in (x0, z0, x1, z1))
    yy = np.full_like(x0, y)
    a = np.stack([x0, yy, z0], -1)
    b = np.stack([x1, yy, z0], -1)
    c = np.stack([x1, yy, z1], -1)
    d = np.stack([x0, yy, z1], -1)
    t1 = np.stack([a, c, b], 1)
    t2 = np.stack([a, d, c], 1)
    verts = np.empty((2 * len(x0), 3, 3))
    verts[0::2] = t1
    verts[1::2] = t2
    normals = np.zeros_like(verts)
    normals[..., 1] = 1.0
    return verts.astype(np.float32), normals.astype(np.float32)


def soup(n: int, seed: int, sigma_median: float = 0.004, sigma_log_sd: float = 0.5):
    """n triangles: centroids U[0,1]^3, vertices c + s*N(0,I), s log-normal."""
    rc = Rng(seed ^ 0x1)
    rs = Rng(seed ^ 0x2)
    rv = Rng(seed ^ 0x3)
    c = rc.uniform(3 * n).reshape(n, 3)
    s = np.exp(math.log(sigma_median) + sigma_log_sd * rs.normal(n))
    off = rv.normal(9 * n).reshape(n, 3, 3)
    verts = (c[:, None, :] + s[:, None, None] * off).astype(np.float32)
    return verts, face_normals(verts)


def _value_noise(p: np.ndarray, seed: int) -> np.ndarray:
    """Trilinear value noise in [-1,1] with smoothstep fade; lattice values hashed by SplitMix64."""
    i = np.floor(p)
    f = p - i
    w = f * f * (3.0 - 2.0 * f)
    ii = i.astype(np.int64)
    lo = ii.min(axis=0)
    hi = ii.max(axis=0) + 1
    ext = hi - lo + 1
    gx, gy, gz = np.meshgrid(*(np.arange(lo[k], hi[k] + 1) for k in range(3)), indexing="ij")
    key = ((gx & 0x1FFFFF).astype(np.uint64)
           | ((gy & 0x1FFFFF).astype(np.uint64) << np.uint64(21))
           | ((gz & 0x1FFFFF).astype(np.uint64) << np.uint64(42)))
    with np.errstate(over="ignore"):
        key = key ^ (np.uint64(seed & 0xFFFFFFFFFFFFFFFF) * _GOLDEN)
    lat = (mix64(key) >> np.uint64(40)).astype(np.float64) * (2.0 / (1 << 24)) - 1.0
    lat = lat.reshape(ext[0], ext[1], ext[2])
    rel = ii - lo
    out = np.zeros(len(p))
    for dx in (0, 1):
        wx = w[:, 0] if dx else 1.0 - w[:, 0]
        for dy in (0, 1):
            wy = w[:, 1] if dy else 1.0 - w[:, 1]
            for dz in (0, 1):
                wz = w[:, 2] if dz else 1.0 - w[:, 2]
                out += wx * wy * wz * lat[rel[:, 0] + dx, rel[:, 1] + dy, rel[:, 2] + dz]
    return out


def displaced_sphere(nu: int, nv: int, seed: int, amplitude: float = 0.15, octaves: int = 6):
    """UV sphere (nu around, nv pole to pole, 2*nu*nv triangles) displaced radially by
    r = 1 + amplitude * sum_o 0.5**o * noise(2**o * u).  Returns verts, normals and the
    per-triangle mean radius (used to band materials)."""
    theta = np.linspace(0.0, math.pi, nv + 1)               # from +y pole
    phi = np.arange(nu) * (2.0 * math.pi / nu)
    th, ph = np.meshgrid(theta, phi, indexing="ij")          # [nv+1, nu]
    u = np.stack([np.sin(th) * np.cos(ph), np.cos(th), np.sin(th) * np.sin(ph)], -1)
    u[0, :] = (0.0, 1.0, 0.0)
    u[-1, :] = (0.0, -1.0, 0.0)
    flat = u.reshape(-1, 3)
    r = np.ones(len(flat))
    for o in range(octaves):
        r += amplitude * (0.5 ** o) * _value_noise(flat * (2.0 ** o) + 0.5, seed + 101 * o)
    r = r.reshape(nv + 1, nu)
    P = u * r[..., None]
    # smooth normals from central differences on the grid
    dphi = np.roll(P, -1, axis=1) - np.roll(P, 1, axis=1)
    Pu = np.concatenate([P[:1], P, P[-1:]], 0)
    dth = Pu[2:] - Pu[:-2]
    n = np.cross(dphi, dth)
    ln = np.linalg.norm(n, axis=-1, keepdims=True)
    ok = ln[..., 0] > 1e-12
    n = np.where(ok[..., None], n / np.where(ok[..., None], ln, 1.0), u)
    flip = np.sum(n * u, axis=-1) < 0
    n[flip] = -n[flip]
    n[0, :] = (0.0, 1.0, 0.0)
    n[-1, :] = (0.0, -1.0, 0.0)
    i = np.arange(nv)[:, None]
    j = np.arange(nu)[None, :]
    jn = (j + 1) % nu
    A = (i, j)
    B = (i, jn)
    C = (i + 1, j)
    D = (i + 1, jn)

    def g(arr, idx):
        return arr[idx[0], idx[1]].reshape(-1, 3)
    ntri = 2 * nu * nv
    verts = np.empty((ntri, 3, 3))
    norms = np.empty((ntri, 3, 3))
    verts[0::2] = np.stack([g(P, A), g(P, C), g(P, D)], 1)
    verts[1::2] = np.stack([g(P, A), g(P, D), g(P, B)], 1)
    norms[0::2] = np.stack([g(n, A), g(n, C), g(n, D)], 1)
    norms[1::2] = np.stack([g(n, A), g(n, D), g(n, B)], 1)
    rt = np.empty(ntri)
    rr = r
    rt[0::2] = (rr[A] + rr[C] + rr[D]).ravel() / 3.0
    rt[1::2] = (rr[A] + rr[D] + rr[B]).ravel() / 3.0
    return verts.astype(np.float32), norms.astype(np.float32), rt


def write_obj(path: str, verts: np.ndarray, normals: np.ndarray) -> None:
    """Writes triangles as a Wavefront OBJ file: one `v` and one `vn` per corner
    (9 significant digits, so float32 values read back exactly) and faces
    `f a//a b//b c//c`."""
    v = np.asarray(verts, np.float32).reshape(-1, 3)
    n = np.asarray(normals, np.float32).reshape(-1, 3)
    with open(path, "w") as f:
        for x, y, z in v:
            f.write(f"v {x:.9g} {y:.9g} {z:.9g}\n")
        for x, y, z in n:
            f.write(f"vn {x:.9g} {y:.9g} {z:.9g}\n")
        for t in range(len(v) // 3):
            a, b, c = 3 * t + 1, 3 * t + 2, 3 * t + 3
            f.write(f"f {a}//{a} {b}//{b} {c}//{c}\n")


def _band(values: np.ndarray, bands: int) -> np.ndarray:
    lo, hi = float(values.min()), float(values.max())
    b = np.floor((values - lo) / max(hi - lo, 1e-30) * bands)
    return np.clip(b, 0, bands - 1).astype(np.int32)


def _random_materials(rng: Rng, m: int, reflective: bool, refl_lo: float, refl_hi: float,
                      kd_lo=0.1, kd_hi=0.9, s_lo=4.0, s_hi=128.0) -> np.ndarray:
    mats = np.zeros(m, MATERIAL_DTYPE)
    kd = kd_lo + (kd_hi - kd_lo) * rng.uniform(3 * m).reshape(m, 3)
    ks = kd_lo + (kd_hi - kd_lo) * rng.uniform(3 * m).reshape(m, 3)
    s = np.rint(s_lo + (s_hi - s_lo) * rng.uniform(m))
    rf = refl_lo + (refl_hi - refl_lo) * rng.uniform(m)
    mats["kd"][:, :3] = kd
    mats["kd"][:, 3] = 1.0
    mats["ks"][:, :3] = ks
    mats["ks"][:, 3] = 1.0
    mats["shininess"] = s
    mats["reflectivity"] = rf if reflective else 0.0
    mats["reflective"] = 1 if reflective else 0
    return mats


def _lights_on_sphere(rng: Rng, k: int, radius: float, center, min_y: float = 0.2) -> np.ndarray:
    d = rng.normal(3 * k).reshape(k, 3)
    d[:, 1] = np.abs(d[:, 1]) + min_y
    d[:, 2] = np.abs(d[:, 2])                    # toward +z (camera side)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return make_lights(np.asarray(center) + radius * d)


# ----------------------------------------------------------------- workloads

def config_a(width: int = 512, height: int = 512) -> SceneData:
    """Icosphere L4 (5,120 tris) over a 64x64-cell plane (8,192 tris); 1 light; depth 1."""
    sv, sn = icosphere(4)
    pv, pn = quad_grid(64, 5.0, -1.0)
    verts = np.concatenate([sv, pv])
    normals = np.concatenate([sn, pn])
    mat = np.concatenate([np.zeros(len(sv), np.int32), np.ones(len(pv), np.int32)])
    mats = np.stack([make_material((0.7, 0.7, 0.7), (0.3, 0.3, 0.3), 32, 0.2, True),
                     make_material((0.7, 0.7, 0.7), (0.0, 0.0, 0.0), 32, 0.0, False)])
    lights = make_lights([(2.0, 4.0, 3.0)])
    cam = make_camera((0.0, 1.0, 4.0), (0.0, 0.0, 0.0), 60.0, width, height)
    return SceneData(verts, normals, mat, mats, lights, cam, 1, "A: icosphere+plane")


def config_b(width: int = 1920, height: int = 1080, n: int = 1_000_000) -> SceneData:
    """1M-triangle soup in the unit cube, 2 lights, reflective material, depth 3."""
    seed = BASE_SEED + 1
    verts, normals = soup(n, seed)
    mat = np.zeros(n, np.int32)
    mats = np.stack([make_material((0.8, 0.6, 0.4), (0.5, 0.5, 0.5), 24, 0.3, True)])
    lights = make_lights([(2.0, 3.0, 3.0), (-1.0, 2.0, 2.0)])
    cam = make_camera((0.5, 0.5, 2.6), (0.5, 0.5, 0.5), 60.0, width, height)
    return SceneData(verts, normals, mat, mats, lights, cam, 3, "B: 1M soup")


def config_c(width: int = 3840, height: int = 2160, nu: int = 2048, nv: int = 1024) -> SceneData:
    """Displaced UV sphere (2*2048*1024 = 4,194,304 tris), 16 banded materials, 4 lights, depth 3."""
    seed = BASE_SEED + 2
    verts, normals, rt = displaced_sphere(nu, nv, seed)
    mat = _band(rt, 16)
    rng = Rng(seed ^ 0x10)
    mats = _random_materials(rng, 16, True, 0.1, 0.5)
    lights = _lights_on_sphere(Rng(seed ^ 0x20), 4, 4.0, (0.0, 0.0, 0.0))
    cam = make_camera((0.0, 0.4, 3.3), (0.0, 0.0, 0.0), 60.0, width, height)
    return SceneData(verts, normals, mat, mats, lights, cam, 3, "C: displaced sphere 4M")


def config_c_dense(width: int = 3840, height: int = 2160, nu: int = 2048, nv: int = 1024,
                   scale: float = 4.0) -> SceneData:
    """Config C's mesh (same 4,194,304 triangles, materials and lights) scaled by
    4 about the origin and seen from close by, so that most primary rays hit.
    At unit scale the reference's absolute |det| < 1e-5 test (src/Triangle.cpp:65)
    rejects every unit-length ray against about 89% of the triangles (their
    doubled area is below 1e-5), so even a camera the sphere fills gets about 3%
    hits; scaled by 4, 95% of the triangles can be hit and about 88% of the
    frame's primary rays hit, each spawning four shadow rays and, on reflective
    materials, a reflection."""
    c = config_c(width, height, nu, nv)
    verts = (c.verts.astype(np.float64) * scale).astype(np.float32)
    lights = c.lights.copy()
    lights["position"][:, :3] = (lights["position"][:, :3].astype(np.float64) * scale).astype(np.float32)
    u = np.array([0.3, 0.45, 1.0])
    u /= np.linalg.norm(u)
    cam = make_camera(1.45 * scale * u, (0.0, 0.0, 0.0), 60.0, width, height)
    return SceneData(verts, c.normals, c.mat, c.materials, lights, cam, c.depth,
                     "C dense: displaced sphere 4M scaled x4, close camera")


def config_d(n_scenes: int = 64, scale: float = 1.0):
    """64 soups, N log-uniform in [10k, 500k], resolution from {256,512,768,1024}^2,
    1-4 lights, depth 2.  ``scale`` shrinks triangle counts and resolutions for tests."""
    seed = BASE_SEED + 3
    rng = Rng(seed)
    out = []
    for k in range(n_scenes):
        u = rng.uniform(4)
        n = int(round(math.exp(math.log(1e4) + u[0] * (math.log(5e5) - math.log(1e4)))))
        res = (256, 512, 768, 1024)[min(int(u[1] * 4), 3)]
        nl = 1 + min(int(u[2] * 4), 3)
        n = max(16, int(n * scale))
        res = max(16, int(res * math.sqrt(scale)) // 8 * 8)
        sseed = seed + 1000 * (k + 1)
        verts, normals = soup(n, sseed, sigma_median=0.004 * math.sqrt(1e6 / n))
        mats = _random_materials(Rng(sseed ^ 0x10), 1, True, 0.2, 0.4, kd_lo=0.3, s_lo=8, s_hi=64)
        lights = _lights_on_sphere(Rng(sseed ^ 0x20), nl, 3.0, (0.5, 0.5, 0.5))
        d = Rng(sseed ^ 0x30).normal(3)
        d[2] = abs(d[2]) + 0.5
        d /= np.linalg.norm(d)
        dist = 2.1 + 0.5 * u[3]
        eye = np.array([0.5, 0.5, 0.5]) + dist * d
        cam = make_camera(eye, (0.5, 0.5, 0.5), 60.0, res, res)
        out.append(SceneData(verts, normals, np.zeros(n, np.int32), mats, lights, cam, 2,
                             f"D[{k}]: soup {n}", {"seed": sseed}))
    return out


def config_e(width: int = 1024, height: int = 1024, nu: int = 1000, nv: int = 500,
             n_materials: int = 256, n_lights: int = 8) -> SceneData:
    """Displaced UV sphere (1,000,000 tris), 256 materials, 8 lights, depth 2."""
    seed = BASE_SEED + 4
    verts, normals, rt = displaced_sphere(nu, nv, seed)
    rng = Rng(seed ^ 0x40)
    # material slots by noise band, then shuffled so neighbouring bands differ
    band = _band(rt, n_materials)
    perm = np.argsort(rng.uniform(n_materials), kind="stable").astype(np.int32)
    mat = perm[band]
    mats = _random_materials(Rng(seed ^ 0x10), n_materials, True, 0.1, 0.4)
    lights = _lights_on_sphere(Rng(seed ^ 0x20), n_lights, 4.0, (0.0, 0.0, 0.0))
    lights["color"][:, :3] = 0.25 + 0.25 * Rng(seed ^ 0x21).uniform(3 * n_lights).reshape(-1, 3)
    cam = make_camera((0.0, 0.3, 3.2), (0.0, 0.0, 0.0), 60.0, width, height)
    return SceneData(verts, normals, mat, mats, lights, cam, 2, "E: adjoint sphere 1M")


def affine(axis, angle: float, scale, translate) -> np.ndarray:
    """Column-major float32 [16] matrix T * R(axis, angle) * S(scale) (glm layout)."""
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    K = np.array([[0.0, -a[2], a[1]], [a[2], 0.0, -a[0]], [-a[1], a[0], 0.0]])
    R = np.eye(3) + math.sin(angle) * K + (1.0 - math.cos(angle)) * (K @ K)
    M = np.eye(4)
    M[:3, :3] = R @ np.diag(np.asarray(scale, np.float64))
    M[:3, 3] = translate
    return M.T.ravel().astype(np.float32)          # column c = M[:, c]


def config_f(width: int = 3840, height: int = 2160, instances: int = 64, nu: int = 256, nv: int = 128,
             rotate: bool = True) -> SceneData:
    """Instanced meshes: one displaced UV sphere (2*nu*nv triangles, local space)
    placed `instances` times on a grid through per-triangle transformations
    (random rotation, non-uniform scale in [0.42, 1.3], one mirrored instance),
    over a world-space ground plane; 9 materials, 2 lights, depth 2.  Defaults:
    64 x 65,536 = 4,194,304 instanced triangles + 2,048 plane triangles.

    rotate=False keeps every matrix axis-aligned (quarter turns about a
    coordinate axis, exact 0/+-1 entries): the only kind the reference's BVH
    bounds correctly, since its Box::transformBoundingBox maps just the min and
    max corners of a local box (src/Box.cu:84-137)."""
    seed = BASE_SEED + 5
    sv, sn, _ = displaced_sphere(nu, nv, seed)
    rng = Rng(seed ^ 0x50)
    cols = int(math.ceil(math.sqrt(instances)))
    rows = int(math.ceil(instances / cols))
    spacing = 2.6
    xforms = []
    for k in range(instances):
        u = rng.uniform(8)
        axis = rng.normal(3)
        s = (0.6 + 0.4 * u[0]) * (0.7 + 0.6 * u[1:4])
        if k == 1:
            s[0] = -s[0]                              # mirrored: negative determinant
        x = (k % cols - 0.5 * (cols - 1)) * spacing
        z = -(k // cols) * spacing
        if rotate:
            xforms.append(affine(axis, 2.0 * math.pi * u[4], s, (x, 0.0, z)))
        else:
            ax = np.zeros(3)
            ax[min(int(u[5] * 3), 2)] = 1.0
            m = affine(ax, 0.5 * math.pi * min(int(u[4] * 4), 3), np.ones(3), (0.0, 0.0, 0.0)).reshape(4, 4)
            m = np.rint(m).astype(np.float64)         # exact quarter turn
            m[:3, :3] *= s[:, None]                   # column-major: row c scales column c, i.e. R * S
            m[3, :3] = (x, 0.0, z)
            xforms.append(m.ravel().astype(np.float32))
    xforms = np.stack(xforms)
    extent = max(cols, rows) * spacing
    pv, pn = quad_grid(32, 0.75 * extent + 2.0, -1.4)
    pv[..., 2] -= 0.5 * (rows - 1) * spacing
    n_inst = len(sv)
    verts = np.concatenate([np.tile(sv, (instances, 1, 1)), pv])
    normals = np.concatenate([np.tile(sn, (instances, 1, 1)), pn])
    xform_index = np.concatenate([np.repeat(np.arange(instances, dtype=np.int32), n_inst),
                                  np.full(len(pv), -1, np.int32)])
    mat = np.concatenate([np.repeat(np.arange(instances, dtype=np.int32) % 8, n_inst),
                          np.full(len(pv), 8, np.int32)])
    mats = np.concatenate([_random_materials(Rng(seed ^ 0x10), 8, True, 0.1, 0.4),
                           make_material((0.6, 0.6, 0.6), (0.1, 0.1, 0.1), 8, 0.0, False)[None]])
    zc = -0.5 * (rows - 1) * spacing
    lights = make_lights([(0.3 * extent, 0.8 * extent + 2.0, zc + 0.6 * extent),
                          (-0.4 * extent, 0.6 * extent + 2.0, zc + 0.2 * extent)])
    cam = make_camera((0.0, 0.3 * extent + 1.0, zc + 0.55 * extent + 1.5), (0.0, -0.4, zc + 0.1 * extent), 60.0,
                      width, height)
    return SceneData(verts, normals, mat, mats, lights, cam, 2, "F: instanced spheres",
                     {"instances": instances, "per_instance": n_inst}, xforms, xform_index)


def glass(width: int = 512, height: int = 384, level: int = 3, depth: int = 3) -> SceneData:
    """Transparent materials: a glass icosphere (refractive index 1.5), a sphere
    that is both reflective and transparent (water, 1.333; its hits spawn two
    rays), a mirror sphere, over a diffuse plane; air outside (1.0002, the
    reference driver's ambient index, src/Material.cpp:25 and src/main.cpp:364)."""
    a, an = icosphere(level, 0.7, (-1.2, 0.0, 0.0))
    b, bn = icosphere(level, 0.7, (0.3, 0.0, -0.6))
    c, cn = icosphere(level, 0.5, (1.4, -0.2, 0.4))
    pv, pn = quad_grid(24, 4.0, -0.7)
    verts = np.concatenate([a, b, c, pv])
    normals = np.concatenate([an, bn, cn, pn])
    mat = np.concatenate([np.full(len(a), 0, np.int32), np.full(len(b), 1, np.int32),
                          np.full(len(c), 2, np.int32), np.full(len(pv), 3, np.int32)])
    mats = np.stack([
        make_material((0.05, 0.05, 0.08), (0.6, 0.6, 0.6), 96, 0.0, False, True, 1.5),
        make_material((0.05, 0.1, 0.15), (0.5, 0.5, 0.5), 48, 0.35, True, True, 1.333),
        make_material((0.2, 0.2, 0.2), (0.8, 0.8, 0.8), 64, 0.6, True),
        make_material((0.6, 0.55, 0.5), (0.1, 0.1, 0.1), 8, 0.0, False)])
    lights = make_lights([(2.0, 4.0, 3.0), (-3.0, 3.0, 1.0)], [(0.8, 0.8, 0.8), (0.5, 0.5, 0.6)])
    cam = make_camera((0.0, 0.9, 3.6), (0.0, -0.1, 0.0), 55.0, width, height)
    s = SceneData(verts, normals, mat, mats, lights, cam, depth, "glass spheres")
    s.ambient_index = 1.0002
    return s


def translation(x: float, y: float, z: float) -> np.ndarray:
    """glm::translate(glm::mat4(1.0f), vec3(x, y, z)) as a column-major float32 [16]."""
    m = np.eye(4, dtype=np.float32)
    m[3, :3] = (x, y, z)          # column 3 of the column-major matrix
    return m.ravel()


def driver_scene(verts: np.ndarray, normals: np.ndarray, width: int = 1280, height: int = 720) -> SceneData:
    """The reference driver's scene (src/main.cpp:330-411, rendered at :454-481):
    one OBJ mesh (the driver loads objmodels/bunny.obj) placed by
    glm::translate(2, -2, -2) through Scene::addTransformation; material
    kd (0.87, 0.3, 0.5, 0), ks 1, shininess 120, neither reflective nor
    transparent; lights at (20, 10, 10) and (-20, 10, 10), colour 1; camera at
    (2, 0, 5) looking down -z with unit right and up vectors (no field of view:
    90 degrees both ways), 1280 x 720; air outside (1.0002); the trace depth
    RayTracer and the renderers default to (4, src/RayTracer.h:36)."""
    n = len(verts)
    mat = make_material((0.87, 0.3, 0.5, 0.0), (1.0, 1.0, 1.0, 1.0), 120, 0.4, False, False, 1.33)
    lights = make_lights([(20.0, 10.0, 10.0), (-20.0, 10.0, 10.0)])
    cam = np.zeros((), CAMERA_DTYPE)
    cam["position"] = (2.0, 0.0, 5.0)
    cam["view"] = (0.0, 0.0, -1.0)
    cam["right"] = (1.0, 0.0, 0.0)
    cam["up"] = (0.0, 1.0, 0.0)
    cam["width"] = width
    cam["height"] = height
    s = SceneData(np.ascontiguousarray(verts, np.float32), np.ascontiguousarray(normals, np.float32),
                  np.zeros(n, np.int32), mat[None], lights, cam, 4, "reference driver scene (bunny)",
                  {}, translation(2.0, -2.0, -2.0)[None], np.zeros(n, np.int32))
    s.ambient_index = 1.0002
    return s


def write_indexed_obj(path: str, positions: np.ndarray, faces: np.ndarray) -> None:
    """Writes `v` positions (9 significant digits: float32 values read back
    exactly) and 1-based triangle `f` lines, no normals."""
    p = np.asarray(positions, np.float32).reshape(-1, 3)
    f = np.asarray(faces, np.int64).reshape(-1, 3) + 1
    with open(path, "w") as out:
        out.writelines(f"v {x:.9g} {y:.9g} {z:.9g}\n" for x, y, z in p)
        out.writelines(f"f {a} {b} {c}\n" for a, b, c in f)


def perturb_params(scene: SceneData, seed: int, rel: float = 0.10):
    """Materials and lights with every parameter scaled by a seeded factor in [1-rel, 1+rel]."""
    rng = Rng(seed)
    mats = scene.materials.copy()
    lights = scene.lights.copy()
    m, n_l = len(mats), len(lights)
    f = 1.0 + rel * (2.0 * rng.uniform(7 * m + 6 * n_l) - 1.0)
    fm = f[:7 * m].reshape(m, 7)
    fl = f[7 * m:].reshape(n_l, 6)
    mats["kd"][:, :3] *= fm[:, 0:3]
    mats["ks"][:, :3] *= fm[:, 3:6]
    mats["shininess"] *= fm[:, 6]
    lights["position"][:, :3] *= fl[:, 0:3]
    lights["color"][:, :3] *= fl[:, 3:6]
    return mats, lights
