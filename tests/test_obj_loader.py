"""The OBJ/MTL reader behind load_model (renderer.cpp:407-603), on the
reference's own assets (assets/CornellBox-Glossy.obj + .mtl, assets/suzanne.obj)
against an independent parse in Python, and on small files written here.
CPU only: vct_obj_open / vct_obj_read need no GPU.  The asset tests are skipped
where the reference tree is absent."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from voxel_cone_tracing_amd import _lib as L

ASSETS = Path("/root/reference/assets")
need_assets = pytest.mark.skipif(not (ASSETS / "CornellBox-Glossy.obj").exists(), reason="reference assets absent")


def py_parse(path: Path):
    """Fan-triangulated faces in file order with their material names, and
    the MTL values, parsed independently of the C++ reader."""
    V, N, tris, mtl = [], [], [], {}
    cur = None
    for line in path.read_text().splitlines():
        t = line.split()
        if not t:
            continue
        if t[0] == "v":
            V.append([float(x) for x in t[1:4]])
        elif t[0] == "vn":
            N.append([float(x) for x in t[1:4]])
        elif t[0] == "mtllib":
            mtl.update(py_mtl(path.parent / t[1]))
        elif t[0] == "usemtl":
            cur = t[1]
        elif t[0] == "f":
            idx = []
            for c in t[1:]:
                p = c.split("/")
                vi = int(p[0])
                ni = int(p[2]) if len(p) > 2 and p[2] else 0
                idx.append((vi - 1 if vi > 0 else len(V) + vi, ni - 1 if ni > 0 else (len(N) + ni if ni < 0 else -1)))
            for k in range(1, len(idx) - 1):
                tri = [idx[0], idx[k], idx[k + 1]]
                tris.append(([V[a] for a, _ in tri], [N[b] if b >= 0 else None for _, b in tri], cur))
    return tris, mtl


def py_mtl(path: Path):
    out, cur = {}, None
    for line in path.read_text().splitlines():
        t = line.split("#")[0].split()
        if not t:
            continue
        if t[0] == "newmtl":
            cur = t[1]
            out[cur] = {}
        elif cur is not None:
            out[cur][t[0]] = [float(x) for x in t[1:]]
    return out


def reference_dims(pos: np.ndarray) -> np.ndarray:
    """renderer.cpp:457-486, 517: max starts at 0, and min.z is updated when a
    coordinate is below min.y (the reference's comparison)."""
    mn = [3.402823466e38] * 3
    mx = [0.0] * 3
    for p in pos.reshape(-1, 3).astype(np.float32):
        x, y, z = float(p[0]), float(p[1]), float(p[2])
        if x > mx[0]:
            mx[0] = x
        if x < mn[0]:
            mn[0] = x
        if y > mx[1]:
            mx[1] = y
        if y < mn[1]:
            mn[1] = y
        if z > mx[2]:
            mx[2] = z
        if z < mn[1]:
            mn[2] = z
    return np.float32(mx) - np.float32(mn)


@need_assets
@pytest.mark.parametrize("name", ["CornellBox-Glossy.obj", "suzanne.obj"])
def test_reference_assets_parse_like_an_independent_reader(name):
    pos, nrm, tm, mats, dims = L.read_obj(ASSETS / name)
    tris, mtl = py_parse(ASSETS / name)
    assert pos.shape[0] == len(tris) > 0
    want = np.array([t[0] for t in tris], np.float32).reshape(-1, 9)
    assert np.allclose(pos, want, rtol=1e-6, atol=0)
    for k, (_, ns, _) in enumerate(tris[:2000]):
        if all(n is not None for n in ns):
            assert np.allclose(nrm[k], np.array(ns, np.float32).reshape(9), rtol=1e-6, atol=1e-7)
    names = list(mtl.keys())
    assert len(mats) == len(names)
    for m, nm in zip(mats, names):
        vals = mtl[nm]
        for key, field in [("Kd", "diffuse"), ("Ks", "specular"), ("Ke", "emission"), ("Ka", "ambient"),
                           ("Tf", "transmittance")]:
            if key in vals:
                assert np.allclose([getattr(m, field)[i] for i in range(3)], vals[key][:3], rtol=1e-6), (nm, key)
        if "Ns" in vals:
            assert m.shininess == pytest.approx(vals["Ns"][0], rel=1e-6)
        if "Ni" in vals:
            assert m.ior == pytest.approx(vals["Ni"][0], rel=1e-6)
        if "illum" in vals:
            assert m.illum == int(vals["illum"][0])
    # per-triangle material: index of the face's usemtl name, -1 without one
    want_tm = [names.index(t[2]) if t[2] in names else -1 for t in tris]
    assert np.array_equal(tm, np.array(want_tm, np.int32))
    assert np.array_equal(dims, reference_dims(pos))


def test_fan_triangulation_negative_indices_and_missing_normals(tmp_path):
    (tmp_path / "m.mtl").write_text("newmtl a\nKd 0.1 0.2 0.3\nNs 7\nd 0.5\nillum 4\nnewmtl b\nKe 1 1 1\n")
    (tmp_path / "q.obj").write_text(
        "mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvn 0 0 1\n"
        "f 1//1 2//1 3//1 4//1\nusemtl b\nf -4 -2 -1\nusemtl nothere\nf 1 2 3\n")
    pos, nrm, tm, mats, dims = L.read_obj(tmp_path / "q.obj")
    assert pos.shape == (4, 9)
    assert np.array_equal(pos[0], [0, 0, 0, 1, 0, 0, 1, 1, 0])
    assert np.array_equal(pos[1], [0, 0, 0, 1, 1, 0, 0, 1, 0])
    assert np.array_equal(pos[2], [0, 0, 0, 1, 1, 0, 0, 1, 0])  # -4 -2 -1 = 1 3 4
    assert np.array_equal(nrm[0], [0, 0, 1] * 3)
    assert np.array_equal(nrm[2], [0, 0, 1] * 3)  # face normal where the file has none
    assert list(tm) == [-1, -1, 1, -1]
    assert mats[0].shininess == 7 and mats[0].dissolve == 0.5 and mats[0].illum == 4
    assert [mats[0].diffuse[i] for i in range(3)] == pytest.approx([0.1, 0.2, 0.3])
    assert [mats[1].emission[i] for i in range(3)] == [1, 1, 1]
    assert mats[1].shininess == 1 and mats[1].ior == 1 and mats[1].dissolve == 1  # defaults
    assert np.array_equal(dims, reference_dims(pos))


def test_missing_file_is_reported():
    with pytest.raises(L.VctError):
        L.read_obj("/nonexistent/x.obj")
