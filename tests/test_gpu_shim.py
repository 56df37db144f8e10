"""The reference-compatible C++ API (include/eos/eos.h) driven by a C++ program
the way the reference's driver does (src/main.cpp:355-481)."""
import os
import subprocess

import numpy as np
import pytest

from eos_amd import scenes
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "eos_amd", "examples", "render_scene")


def _write_scene(path, s):
    with open(path, "wb") as f:
        np.array([s.n_tri, len(s.materials), len(s.lights), s.width, s.height, s.depth], np.int32).tofile(f)
        np.ascontiguousarray(s.verts, np.float32).tofile(f)
        np.ascontiguousarray(s.normals, np.float32).tofile(f)
        np.ascontiguousarray(s.mat, np.int32).tofile(f)
        np.ascontiguousarray(s.materials).tofile(f)
        np.ascontiguousarray(s.lights).tofile(f)
        np.concatenate([s.camera["position"], s.camera["view"], s.camera["right"],
                        s.camera["up"]]).astype(np.float32).tofile(f)


def _shim_scene(name):
    if name == "config_a":
        return scenes.config_a(80, 60)
    if name == "soup":
        return scenes.config_b(72, 48, n=8000)
    s = scenes.glass(80, 60, level=2)           # transparent materials through Material::setTransparent
    s.ambient_index = 1.0                        # the C++ Scene's default
    return s


@pytest.mark.parametrize("name", ["config_a", "soup", "glass"])
def test_cpp_api_renders_like_the_c_abi(tmp_path, name):
    from eos_amd import _lib as E
    s = _shim_scene(name)
    sc = tmp_path / "scene.bin"
    _write_scene(sc, s)
    o1, o2 = tmp_path / "multi.rgba", tmp_path / "single.rgba"
    print(f"running {DEMO} on {name}", flush=True)
    aa = 2 if name == "config_a" else 0           # the RayTracer route with renderWithAA
    r = subprocess.run([DEMO, str(sc), str(o1), str(o2)], capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, EOS_QUIET="1", EOS_AA=str(aa)))
    assert r.returncode == 0, r.stderr
    multi = np.fromfile(o1, np.uint8).reshape(s.height, s.width, 4)
    single = np.fromfile(o2, np.uint8).reshape(s.height, s.width, 4)
    g = E.Scene(s, leaf_max=4)
    ref = g.render()["rgba"]
    assert np.array_equal(multi, ref)
    assert np.array_equal(single, ref)
    rt = np.fromfile(str(o2) + ".rt", np.uint8).reshape(s.height, s.width, 4)
    assert np.array_equal(rt, g.render(aa_samples=aa)["rgba"])
    if aa:
        assert np.abs(rt.astype(int) - po.Oracle(s).render_aa(aa)["rgba"].astype(int)).max(-1).__le__(1).mean() \
            >= 0.9999
    orc = po.Oracle(s).render()["rgba"]
    assert (np.abs(multi.astype(int) - orc.astype(int)).max(-1) <= 1).mean() >= 0.9999


def test_cpp_api_triangle_mesh_from_obj(tmp_path):
    """Config A with its sphere read from an OBJ file through TriangleMesh and
    Scene::addTriangleMesh renders like the whole scene through the C ABI."""
    from eos_amd import _lib as E
    s = scenes.config_a(64, 48)
    sphere = s.mat == s.mat[0]
    assert sphere.sum() == 5120 and np.all(np.flatnonzero(sphere) == np.arange(5120))
    rest = scenes.SceneData(s.verts[~sphere], s.normals[~sphere], s.mat[~sphere], s.materials, s.lights, s.camera,
                            s.depth, "plane")
    sc, obj = tmp_path / "plane.bin", tmp_path / "sphere.obj"
    _write_scene(sc, rest)
    scenes.write_obj(str(obj), s.verts[sphere], s.normals[sphere])
    o1, o2 = tmp_path / "multi.rgba", tmp_path / "single.rgba"
    print("running the C++ example with an OBJ mesh", flush=True)
    r = subprocess.run([DEMO, str(sc), str(o1), str(o2), str(obj), str(int(s.mat[0]))], capture_output=True,
                       text=True, timeout=120, env=dict(os.environ, EOS_QUIET="1"))
    assert r.returncode == 0, r.stderr
    assert "has 5120 Triangles" in r.stdout
    # the C++ scene holds the plane first, then the mesh
    order = np.concatenate([np.flatnonzero(~sphere), np.flatnonzero(sphere)])
    whole = scenes.SceneData(s.verts[order], s.normals[order], s.mat[order], s.materials, s.lights, s.camera,
                             s.depth, "reordered")
    ref = E.Scene(whole, leaf_max=4).render()["rgba"]
    multi = np.fromfile(o1, np.uint8).reshape(s.height, s.width, 4)
    single = np.fromfile(o2, np.uint8).reshape(s.height, s.width, 4)
    assert np.array_equal(multi, ref) and np.array_equal(single, ref)


def test_cpp_api_translated_mesh(tmp_path):
    """The reference driver's route for its OBJ model (src/main.cpp:386-403): a
    translation through Scene::addTransformation and the mesh's transformation
    index renders like the C ABI with that per-triangle transformation."""
    from eos_amd import _lib as E
    s = scenes.config_a(64, 48)
    sphere = s.mat == s.mat[0]
    rest = scenes.SceneData(s.verts[~sphere], s.normals[~sphere], s.mat[~sphere], s.materials, s.lights, s.camera,
                            s.depth, "plane")
    sc, obj = tmp_path / "plane.bin", tmp_path / "sphere.obj"
    _write_scene(sc, rest)
    scenes.write_obj(str(obj), s.verts[sphere], s.normals[sphere])
    o1, o2 = tmp_path / "multi.rgba", tmp_path / "single.rgba"
    t = ("0.75", "-0.25", "-0.5")
    print("running the C++ example with a translated OBJ mesh", flush=True)
    r = subprocess.run([DEMO, str(sc), str(o1), str(o2), str(obj), str(int(s.mat[0])), *t], capture_output=True,
                       text=True, timeout=120, env=dict(os.environ, EOS_QUIET="1"))
    assert r.returncode == 0, r.stderr
    order = np.concatenate([np.flatnonzero(~sphere), np.flatnonzero(sphere)])
    m = np.eye(4, dtype=np.float32)
    m[3, :3] = [float(x) for x in t]               # column-major: column 3 holds the translation
    xi = np.concatenate([np.full((~sphere).sum(), -1, np.int32), np.zeros(sphere.sum(), np.int32)])
    whole = scenes.SceneData(s.verts[order], s.normals[order], s.mat[order], s.materials, s.lights, s.camera,
                             s.depth, "translated", {}, m.ravel()[None], xi)
    ref = E.Scene(whole, leaf_max=4).render(want_hits=True)
    multi = np.fromfile(o1, np.uint8).reshape(s.height, s.width, 4)
    single = np.fromfile(o2, np.uint8).reshape(s.height, s.width, 4)
    assert np.array_equal(multi, ref["rgba"]) and np.array_equal(single, ref["rgba"])
    # the sphere moved: the oracle with the same transformation agrees
    oo = po.Oracle(whole).render()
    assert (ref["hit"] == oo["hit"]).mean() >= 0.9999
    plain = E.Scene(scenes.SceneData(whole.verts, whole.normals, whole.mat, s.materials, s.lights, s.camera,
                                     s.depth, "unmoved"), leaf_max=4).render()["rgba"]
    assert not np.array_equal(plain, ref["rgba"])


@pytest.mark.parametrize("n_bands", [2, 3, 7])
def test_cpp_split_frame_launcher_renders_each_band(tmp_path, n_bands):
    """rayTrace_MultiMegaKernel(…, numGPUS, gpuID, …) called once per band, as
    the reference's MultiDeviceRenderer threads call it: each band is the
    frame's rows [k·H/n, (k+1)·H/n) at the launcher's traced depth (4), byte for
    byte; rows past n·(H/n) belong to no band, as in the reference."""
    from eos_amd import _lib as E
    s = scenes.config_b(72, 50, n=8000)
    sc = tmp_path / "scene.bin"
    _write_scene(sc, s)
    o1, o2 = tmp_path / "multi.rgba", tmp_path / "single.rgba"
    print(f"running the C++ example with {n_bands} bands", flush=True)
    r = subprocess.run([DEMO, str(sc), str(o1), str(o2)], capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, EOS_QUIET="1", EOS_BANDS=str(n_bands)))
    assert r.returncode == 0, r.stderr
    rows = s.height // n_bands
    bands = np.fromfile(str(o2) + ".bands", np.uint8).reshape(rows * n_bands, s.width, 4)
    g = E.Scene(s, leaf_max=4)
    full = g.render(depth=4)["rgba"]
    assert np.array_equal(bands, full[:rows * n_bands])


def test_cpp_api_reference_driver_scene(tmp_path):
    """The reference driver's own scene built the way its driver builds it
    (src/main.cpp:330-481): the bunny read through TriangleMesh::loadFromFile
    (the committed mesh written back as OBJ), glm::translate(2, -2, -2) through
    Scene::addTransformation and the mesh's transformation index, its material
    and two lights, then MultiDeviceRenderer at 1280 x 720 and the single-GPU
    route.  The image equals the C ABI's and agrees with the oracle; the Scene's
    ray queries (per ray and batched) answer the reference's golden rays as the
    reference did."""
    from eos_amd import _lib as E
    from tests.test_golden import GOLD, bunny_scene
    gold = np.load(GOLD, allow_pickle=False)
    s = bunny_scene()
    m = np.load(os.path.join(ROOT, "tests", "golden", "bunny_mesh.npz"), allow_pickle=False)
    empty = scenes.SceneData(np.zeros((0, 3, 3), np.float32), np.zeros((0, 3, 3), np.float32), np.zeros(0, np.int32),
                             s.materials, s.lights, s.camera, s.depth, "driver scene without its mesh")
    sc, obj = tmp_path / "driver.bin", tmp_path / "bunny.obj"
    _write_scene(sc, empty)
    scenes.write_indexed_obj(str(obj), m["positions"], m["faces"])
    hit = gold["bunny/lin_hit"].astype(bool)
    rng = np.random.default_rng(7)
    sel = np.sort(np.concatenate([np.flatnonzero(hit), rng.choice(np.flatnonzero(~hit), 200, replace=False)]))
    qo = np.concatenate([gold["bunny/ray_o"][sel], gold["bunny/shadow_o"]]).astype(np.float32)
    qd = np.concatenate([gold["bunny/ray_d"][sel], gold["bunny/shadow_d"]]).astype(np.float32)
    qf = tmp_path / "rays.bin"
    with open(qf, "wb") as f:
        np.array([len(qo)], np.int32).tofile(f)
        qo.tofile(f)
        qd.tofile(f)
    o1, o2 = tmp_path / "multi.rgba", tmp_path / "single.rgba"
    print(f"running the C++ example on the driver scene ({s.n_tri} triangles, {len(qo)} query rays)", flush=True)
    r = subprocess.run([DEMO, str(sc), str(o1), str(o2), str(obj), "0", "2", "-2", "-2"], capture_output=True,
                       text=True, timeout=300, env=dict(os.environ, EOS_QUIET="1", EOS_QUERIES=str(qf)))
    assert r.returncode == 0, r.stderr
    assert "has 69666 Triangles" in r.stdout
    W, H = 1280, 720
    multi = np.fromfile(o1, np.uint8).reshape(H, W, 4)
    single = np.fromfile(o2, np.uint8).reshape(H, W, 4)
    ref = E.Scene(s, leaf_max=4).render(want_hits=True)
    assert np.array_equal(multi, ref["rgba"]) and np.array_equal(single, ref["rgba"])
    oo = po.Oracle(s).render()
    agree = (ref["hit"] == oo["hit"]).mean()
    assert agree >= 0.9999, agree
    diff = np.abs(ref["rgba"].astype(int) - oo["rgba"].astype(int)).max(-1)
    assert (diff[ref["hit"] == oo["hit"]] <= 1).all()
    assert (ref["hit"] >= 0).sum() > 10000                   # the bunny is in the frame
    rec = np.dtype([("found", np.int32), ("point", np.float32, 3), ("normal", np.float32, 3), ("mat", np.int32),
                    ("occ", np.int32), ("lfound", np.int32), ("lpoint", np.float32, 3), ("lnormal", np.float32, 3),
                    ("lmat", np.int32)])
    per_ray = np.fromfile(str(qf) + ".out", rec)
    batched = np.fromfile(str(qf) + ".batch", rec)
    assert len(per_ray) == len(qo) and np.array_equal(per_ray.view(np.uint8), batched.view(np.uint8))
    n1 = len(sel)
    want_hit = np.concatenate([hit[sel], gold["bunny/shadow_hit"].astype(bool)])
    assert np.array_equal(per_ray["lfound"] == 1, want_hit)
    assert np.array_equal(per_ray["found"] == 1, want_hit) and np.array_equal(per_ray["occ"] == 1, want_hit)
    f = per_ray["lfound"][:n1] == 1
    assert np.array_equal(per_ray["lpoint"][:n1][f], gold["bunny/lin_point"][sel][f])
    assert np.array_equal(per_ray["lnormal"][:n1][f], gold["bunny/lin_normal"][sel][f])
    assert np.array_equal(per_ray["point"][:n1][f], gold["bunny/lin_point"][sel][f])
    assert (per_ray["mat"][per_ray["found"] == 1] == 0).all()
