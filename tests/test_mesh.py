"""Wavefront OBJ reading (eos_mesh_load_obj), checked without a GPU.

Positions, corner normals and triangle fans follow from the file alone; the
smooth normals of files without `vn` are checked against their closed forms
(flat patch: the face normal; cube corner: (1,1,1)/sqrt(3)).  The reference
generates those through assimp, which is not available: generated normals are
parity unpinned against it.
"""
import os

import numpy as np
import pytest

from eos_amd import _lib as E

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _write(tmp_path, text, name="m.obj"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_corner_formats_fans_and_negative_indices(tmp_path):
    path = _write(tmp_path, """# comment
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
vt 0 0
vn 0 0 1
vn 0 0 -1

f 1 2 3 4
f 1/1 3/1 2/1
f 1//2 2//2 3//2
f -4/1/1 -3/1/1 -2/1/1
""")
    v, n = E.load_obj(path)
    assert v.shape == (5, 3, 3)
    P = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32)
    # quad as a fan from its first corner
    assert np.array_equal(v[0], P[[0, 1, 2]]) and np.array_equal(v[1], P[[0, 2, 3]])
    assert np.array_equal(v[2], P[[0, 2, 1]])
    assert np.array_equal(v[3], P[[0, 1, 2]]) and np.array_equal(v[4], P[[0, 1, 2]])
    assert np.array_equal(n[3], np.tile([0, 0, -1], (3, 1)).astype(np.float32))
    assert np.array_equal(n[4], np.tile([0, 0, 1], (3, 1)).astype(np.float32))
    # corners without vn: smooth normal of a flat patch = its face normal (+z);
    # the wound-backwards face 2 cancels nothing because every face here is flat in z
    for t in (0, 1):
        assert np.allclose(np.abs(n[t]), np.tile([0, 0, 1], (3, 1)), atol=0)


def test_smooth_normals_cube_corner(tmp_path):
    # three faces meeting at the origin corner of a unit cube, outward normals -x, -y, -z
    path = _write(tmp_path, """v 0 0 0
v 1 0 0
v 0 1 0
v 0 0 1
f 1 3 2
f 1 2 4
f 1 4 3
""")
    v, n = E.load_obj(path)
    assert v.shape == (3, 3, 3)
    want = -np.ones(3) / np.sqrt(3.0)
    for t in range(3):
        assert np.allclose(n[t, 0], want, atol=1e-7)          # the shared corner
    # vertex 2 (1,0,0) is used by faces 1 and 2: normals -z and -y
    assert np.allclose(n[0, 2], np.array([0, -1, -1]) / np.sqrt(2.0), atol=1e-7)


def test_errors(tmp_path):
    with pytest.raises(E.EosError) as ei:
        E.load_obj(str(tmp_path / "missing.obj"))
    assert ei.value.status == E.EOS_ERR_INVALID
    bad = _write(tmp_path, "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 9\n", "bad.obj")
    with pytest.raises(E.EosError) as ei:
        E.load_obj(bad)
    assert "bad.obj:4" in str(ei.value)
    two = _write(tmp_path, "v 0 0 0\nv 1 0 0\nf 1 2\n", "two.obj")
    with pytest.raises(E.EosError):
        E.load_obj(two)
    # indices that wrap to a valid int when cast (2^32 + 1 -> 1, -(2^32) + 1) are out of range
    for k, idx in enumerate(("4294967297", "-4294967295", "99999999999999999999", "2147483648")):
        big = _write(tmp_path, f"v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nf 1 2 {idx}\n", f"big{k}.obj")
        with pytest.raises(E.EosError) as ei:
            E.load_obj(big)
        assert ei.value.status == E.EOS_ERR_INVALID, idx
        bign = _write(tmp_path, f"v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nf 1//1 2//1 3//{idx}\n", f"bign{k}.obj")
        with pytest.raises(E.EosError):
            E.load_obj(bign)
    empty = _write(tmp_path, "# nothing\n", "empty.obj")
    v, n = E.load_obj(empty)
    assert v.shape == (0, 3, 3)


def test_round_trip_of_a_written_mesh(tmp_path):
    from eos_amd import scenes
    s = scenes.config_a(8, 8)
    path = str(tmp_path / "scene.obj")
    scenes.write_obj(path, s.verts, s.normals)
    v, n = E.load_obj(path)
    assert np.array_equal(v, s.verts) and np.array_equal(n, s.normals)


REF_MODELS = "/root/reference/objmodels"


@pytest.mark.skipif(not os.path.isdir(REF_MODELS), reason="the reference's OBJ models are not present")
@pytest.mark.parametrize("name", ["bunny.obj", "head.obj", "monkey.obj", "grid.obj", "grid_back.obj",
                                  "triangle_grid.obj"])
def test_reference_models_read_exactly(name):
    """The OBJ models the reference ships (its driver loads bunny.obj,
    src/main.cpp:386-403): every face becomes its fan of triangles, every
    corner the file's position, and every generated normal is a unit vector."""
    path = os.path.join(REF_MODELS, name)
    pos, faces = [], []
    with open(path) as f:
        for line in f:
            t = line.split()
            if not t:
                continue
            if t[0] == "v":
                pos.append([float(x) for x in t[1:4]])
            elif t[0] == "f":
                faces.append([int(c.split("/")[0]) for c in t[1:]])
    P = np.asarray(pos, np.float32)
    want = []
    for c in faces:
        idx = [i - 1 if i > 0 else len(P) + i for i in c]
        for k in range(1, len(idx) - 1):
            want.append(P[[idx[0], idx[k], idx[k + 1]]])
    v, n = E.load_obj(path)
    assert v.shape == (len(want), 3, 3)
    assert np.array_equal(v, np.asarray(want, np.float32).reshape(-1, 3, 3))
    ln = np.linalg.norm(n.astype(np.float64), axis=-1)
    assert np.all((np.abs(ln - 1.0) < 1e-6) | (ln == 0.0))
    if name == "bunny.obj":
        assert len(v) == 69666
