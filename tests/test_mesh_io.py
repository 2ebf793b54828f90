"""Mesh and result files in the reference harness's formats: the .pbf writer
(read back by Read_model_data, bem-bb-fw-HACApK-0.4.2.f90:411-640) and the VTK
writer (Print_result, :793-855)."""
import hashlib
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_pbf_writer_reproduces_reference_sample(H, tmp_path):
    """Writing the sample mesh's nodes, faces and face parameters gives the
    reference's input_sample.pbf byte for byte (its CRLF line ends as LF)."""
    d = np.load(os.path.join(GOLD, "input_sample_mesh.npz"))
    want = json.load(open(os.path.join(GOLD, "input_sample_pbf.json")))
    path = tmp_path / "s.pbf"
    H.write_pbf(path, d["nodes"], d["faces"], d["dble_para"])
    raw = path.read_bytes()
    assert len(raw) == want["bytes"]
    assert hashlib.sha256(raw).hexdigest() == want["sha256"]


def test_pbf_round_trip(H, tmp_path):
    nodes, faces = H.icosphere(3, 2.5, (1.0, -2.0, 0.5))
    rng = np.random.default_rng(4)
    para = rng.uniform(-1e3, 1e3, (len(faces), 2))
    path = tmp_path / "r.pbf"
    H.write_pbf(path, nodes, faces, para)
    n2, f2, p2 = H.read_pbf(str(path))
    assert np.array_equal(f2, faces)
    # 13 significant digits in the file
    assert np.allclose(n2, nodes, rtol=1e-12, atol=1e-15)
    assert np.allclose(p2, para[:, 0], rtol=1e-12)
    H.write_pbf(tmp_path / "none.pbf", nodes, faces)
    n3, f3, p3 = H.read_pbf(str(tmp_path / "none.pbf"))
    assert np.array_equal(f3, faces) and np.all(p3 == 0.0)
    with pytest.raises(H.HmvError):
        H.write_pbf(tmp_path / "bad.pbf", nodes, faces + len(nodes))


def test_vtk_writer(H, tmp_path):
    nodes, faces = H.icosphere(2)
    sol = H.uniform(5, len(faces)) * np.logspace(-6, 2, len(faces))
    path = tmp_path / "out.vtk"
    H.write_vtk(path, nodes, faces, sol)
    lines = path.read_text().splitlines()
    assert lines[:4] == ["# vtk DataFile Version 2.0", "header : this data is results of BEM-BB",
                         "ASCII", "DATASET UNSTRUCTURED_GRID"]
    nn, nf = len(nodes), len(faces)
    assert lines[4] == "POINTS %d float" % nn
    pts = np.array([[float(v) for v in l.split()] for l in lines[5:5 + nn]])
    assert all(len(l) == 50 for l in lines[5:5 + nn])  # 3 x f16.13 with single blanks
    assert np.allclose(pts, nodes, atol=6e-14)
    k = 5 + nn
    assert lines[k] == "CELLS %d %d" % (nf, 4 * nf)
    cells = np.array([[int(v) for v in l.split()] for l in lines[k + 1:k + 1 + nf]])
    assert np.all(cells[:, 0] == 3) and np.array_equal(cells[:, 1:], faces)
    k += 1 + nf
    assert lines[k] == "CELL_TYPES %d" % nf
    assert all(l.split() == ["5"] for l in lines[k + 1:k + 1 + nf])
    k += 1 + nf
    assert lines[k].split() == ["CELL_DATA", str(nf)]
    assert lines[k + 1].split() == ["SCALARS", "solve", "float", "1"]
    assert lines[k + 2].split() == ["LOOKUP_TABLE", "default"]
    vals = lines[k + 3:k + 3 + nf]
    assert all(len(l) == 32 for l in vals)  # f32.28
    assert np.array_equal(np.array([float(v) for v in vals]), sol)
    assert len(lines) == k + 3 + nf
