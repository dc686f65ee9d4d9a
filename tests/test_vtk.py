"""VTK XML particle output of the Fortran shim in the reference's layout
(src/utils/module_vtk.f90, module_vtk_helpers.f90:49-113,
coulomb/module_interaction_specific_types_XYZQ.f90:63-75): per step an
unstructured-grid piece ./vtk/<name>_<step>.000000.vtu, the parallel index
./vtk/<name>_<step>.pvtu, and the time series ./<name>.timeseries.pvd / .visit.
Every binary data array is base64 text of a big-endian 4-byte byte count
followed by base64 text of the big-endian values; this test decodes them with
Python's base64 and numpy and checks them against the inputs.  Runs on the CPU
(the writer needs no GPU context)."""
import base64
import os
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "pepc_amd", "fortran", "pepc_mini_gpu")
DT = {"Float64": ">f8", "Int64": ">i8", "Int32": ">i4"}


def _decode(text, vtype):
    text = text.strip()
    head = base64.b64decode(text[:8])
    nbytes = int(np.frombuffer(head, ">u4")[0])
    body = base64.b64decode(text[8:])
    assert len(body) == nbytes
    return np.frombuffer(body, DT[vtype])


def _arrays(path):
    root = ET.parse(path).getroot()
    assert root.get("type") == "UnstructuredGrid" and root.get("byte_order") == "BigEndian"
    piece = root.find("UnstructuredGrid/Piece")
    out = {}
    for sec in ("Points", "PointData", "Cells", "CellData"):
        for da in piece.find(sec).findall("DataArray"):
            assert da.get("format") == "binary"
            v = _decode(da.text, da.get("type"))
            nc = int(da.get("NumberOfComponents"))
            out[da.get("Name")] = v.reshape(-1, nc) if nc > 1 else v
    return piece, out


def test_particle_series(tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("Fortran frontend not built (amdflang absent)")
    n = 77
    r = np.random.default_rng(7)
    x, y, z = r.random((3, n))
    q = r.normal(size=n)
    with open(tmp_path / "in.bin", "wb") as f:
        np.array([n], np.int64).tofile(f)
        for a in (x, y, z, q):
            a.tofile(f)
    res = subprocess.run([EXE, "--vtk", "in.bin", "parts"], cwd=tmp_path, capture_output=True,
                         text=True, timeout=60)
    assert res.returncode == 0, res.stdout + res.stderr
    for k in range(3):
        piece, a = _arrays(tmp_path / "vtk" / f"parts_{k:06d}.000000.vtu")
        assert piece.get("NumberOfPoints") == str(n) and piece.get("NumberOfCells") == str(n)
        scale = 2.0 if k == 2 else 1.0
        assert np.array_equal(a["xyz"], scale * np.stack([x, y, z], 1))
        assert np.array_equal(a["work"], np.full(n, 1.0 + k))
        assert np.array_equal(a["label"], np.arange(n))
        assert np.array_equal(a["local index"], np.arange(1, n + 1))
        assert np.array_equal(a["processor"], np.zeros(n))
        assert np.array_equal(a["q"], q)
        assert np.array_equal(a["pot"], q * (k + 1))
        assert np.array_equal(a["field"], np.stack([x * (k + 1), -y, z + k], 1))
        assert ("L2 error" in a) == (k == 1)
        if k == 1:
            assert np.array_equal(a["L2 error"], 2.0 * q * 2)
        assert np.array_equal(a["connectivity"], np.arange(n))
        assert np.array_equal(a["offsets"], np.arange(1, n + 1))
        assert np.array_equal(a["types"], np.ones(n))       # VTK_VERTEX
        # parallel index: same array declarations, one piece
        pv = ET.parse(tmp_path / "vtk" / f"parts_{k:06d}.pvtu").getroot()
        assert pv.get("type") == "PUnstructuredGrid"
        names = [d.get("Name") for d in pv.iter("DataArray")]
        assert sorted(names) == sorted(a.keys())
        assert [p.get("Source") for p in pv.iter("Piece")] == [f"parts_{k:06d}.000000.vtu"]
    pvd = ET.parse(tmp_path / "parts.timeseries.pvd").getroot()
    sets = pvd.findall("Collection/DataSet")
    assert [float(d.get("timestep")) for d in sets] == [0.0, 0.5, 1.0]
    assert [d.get("file") for d in sets] == [f"./vtk/parts_{k:06d}.pvtu" for k in range(3)]
    visit = (tmp_path / "parts.timeseries.visit").read_text().split()
    assert visit[:2] == ["!NBLOCKS", "1"]
    assert visit[2:] == [f"./vtk/parts_{k:06d}.000000.vtu" for k in range(3)]


def test_spacecurve_and_tree_writers(tmp_path):
    """vtk_write_spacecurve (module_vtk_helpers.f90:724-764): the particles in
    order joined by one poly-line cell (VTK type 4) tagged with the rank.
    vtk_write_branches / vtk_write_leaves see an unallocated host tree (the
    tree lives on the device) and skip with the reference's message."""
    if not os.path.exists(EXE):
        pytest.skip("Fortran frontend not built (amdflang absent)")
    n = 40
    r = np.random.default_rng(8)
    x, y, z = r.random((3, n))
    with open(tmp_path / "in.bin", "wb") as f:
        np.array([n], np.int64).tofile(f)
        for a in (x, y, z, r.normal(size=n)):
            a.tofile(f)
    res = subprocess.run([EXE, "--vtk", "in.bin", "parts"], cwd=tmp_path, capture_output=True,
                         text=True, timeout=60)
    assert res.returncode == 0, res.stdout + res.stderr
    assert res.stdout.count("tree is not allocated, aborting branch output") == 3
    assert res.stdout.count("tree is not allocated, aborting leaves output") == 3
    for k in range(3):
        piece, a = _arrays(tmp_path / "vtk" / f"spacecurve_{k:06d}.000000.vtu")
        assert piece.get("NumberOfPoints") == str(n) and piece.get("NumberOfCells") == "1"
        assert np.array_equal(a["xyz"], np.stack([x, y, z], 1))
        assert np.array_equal(a["connectivity"], np.arange(n))
        assert a["offsets"].tolist() == [n] and a["types"].tolist() == [4]
        assert a["processor"].tolist() == [0]
    sets = ET.parse(tmp_path / "spacecurve.timeseries.pvd").getroot().findall("Collection/DataSet")
    assert [d.get("file") for d in sets] == [f"./vtk/spacecurve_{k:06d}.pvtu" for k in range(3)]
