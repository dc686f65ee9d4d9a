"""On-disk outputs of a GPU solve through the Fortran frontend path: the
particle checkpoint (utils/module_checkpoint.f90:122-300) written after a
grow + traverse and read back, and the tree writers vtk_write_branches /
vtk_write_leaves (utils/module_vtk_helpers.f90:233-373, 553-723) on the
grown device tree, all against the oracle's solve of the same particles:
records bit-exact (fields to 1e-12 in FP64), voxel corners, levels, leaf and
descendant counts and node charges bit-exact.  After pepc_timber_tree the
leaves writer takes the reference's unallocated-tree path."""
import base64
import os
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from pepc_amd import workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "pepc_amd", "fortran", "pepc_mini_gpu")
REC = np.dtype([("x", "<f8", (3,)), ("work", "<f8"), ("key", "<i8"), ("node_leaf", "<i8"),
                ("label", "<i8"), ("q", "<f8"), ("e", "<f8", (3,)), ("pot", "<f8")])
DT = {"Float64": ">f8", "Int64": ">i8", "Int32": ">i4"}
# voxel corner j: offsets (x, y, z) in units of the node's box size
SHIFT = np.array([[0, 0, 0], [0, 0, 1], [0, 1, 0], [0, 1, 1],
                  [1, 0, 0], [1, 0, 1], [1, 1, 0], [1, 1, 1]], np.float64)

pytestmark = pytest.mark.gpu


def _vtu(path):
    piece = ET.parse(path).getroot().find("UnstructuredGrid/Piece")
    out = {}
    for sec in ("Points", "PointData", "Cells", "CellData"):
        for da in piece.find(sec).findall("DataArray"):
            t = da.text.strip()
            nb = int(np.frombuffer(base64.b64decode(t[:8]), ">u4")[0])
            body = base64.b64decode(t[8:])
            assert len(body) == nb
            v = np.frombuffer(body, DT[da.get("type")])
            nc = int(da.get("NumberOfComponents"))
            out[da.get("Name")] = v.reshape(-1, nc) if nc > 1 else v
    return piece, out


def _expected_boxes(orc, d, t, nodes):
    from oracle import oracle as O
    bmin, bmax, bsize = d.box.as_arrays()
    corners, desc = [], []
    for k in nodes:
        x = orc.key_to_coord(d.box, t.key[k])
        s = bsize / 2.0 ** t.level[k]
        corners.append(x[None, :] + SHIFT * s[None, :])
        # descendants: the nodes below k (keys with k's key as their prefix)
        lv = int(t.level[k])
        deeper = t.level >= lv
        shift = (3 * (t.level[deeper] - lv)).astype(np.uint64)
        desc.append(int(np.sum((t.key[deeper] >> shift) == t.key[k])) - 1)
    return (np.concatenate(corners), np.array(desc, np.int64),
            t.mom[nodes, O.CHARGE], t.mom[nodes, O.ABSCHARGE])


def test_checkpoint_and_tree_output_of_a_gpu_solve(orc, tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("pepc_mini_gpu not built (python -m pepc_amd.build)")
    s = workloads.mini(n=6000, seed=47)
    fin, fout, par = tmp_path / "in.bin", tmp_path / "out.bin", tmp_path / "params"
    with open(fin, "wb") as f:
        np.array([s.n], np.int64).tofile(f)
        for a in (s.x, s.y, s.z, s.q):
            a.astype(np.float64).tofile(f)
    with open(par, "w") as f:
        f.write(f"&calc_force_coulomb\n theta2 = {s.theta2!r}\n eps2 = {s.eps2!r}\n/\n")
        f.write("&walk_para_gpu\n group_size = 16\n precision = 'fp64'\n/\n")
    r = subprocess.run([EXE, "--solve-output", str(par), str(fin), str(fout)], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    d, t, w = orc.solve(s.x, s.y, s.z, s.q, s.theta2, s.eps2, tile_size=16)

    # checkpoint read back: header {n, itime} then the records in key order
    raw = open(fout, "rb").read()
    assert np.frombuffer(raw[:16], "<i8").tolist() == [s.n, 3]
    rec = np.frombuffer(raw[16:], REC)
    assert len(rec) == s.n
    assert np.array_equal(rec["x"], np.stack([d.x, d.y, d.z], 1))
    assert np.array_equal(rec["key"].astype(np.uint64), d.keys)
    assert np.array_equal(rec["label"], d.perm)
    assert np.array_equal(rec["q"], d.q)
    assert np.array_equal(rec["work"], w.work)
    assert np.max(np.abs(rec["e"] - w.e)) <= 1e-12 * np.abs(w.e).max()
    assert np.max(np.abs(rec["pot"] - w.pot)) <= 1e-12 * np.abs(w.pot).max()
    mpi = (tmp_path / "particles" / "mpi" / "particle_000000000003.mpi").read_bytes()
    assert mpi[128:] == rec.tobytes()

    leaf = t.first_child < 0
    # leaves: one voxel per particle, in key order
    lnodes = np.nonzero(leaf)[0]
    lnodes = lnodes[np.argsort(t.first[lnodes], kind="stable")]
    piece, a = _vtu(tmp_path / "vtk" / "leaves_000003.000000.vtu")
    assert piece.get("NumberOfCells") == str(s.n)
    corners, desc, ch, ach = _expected_boxes(orc, d, t, lnodes)
    assert np.array_equal(a["corners"], corners)
    assert np.array_equal(a["level"], t.level[lnodes])
    assert np.array_equal(a["leaves"], np.ones(s.n))
    assert np.array_equal(a["descendants"], np.zeros(s.n))
    assert np.array_equal(a["charge"], ch) and np.array_equal(a["abs_charge"], ach)
    assert np.array_equal(a["connectivity"], np.arange(8 * s.n))
    assert np.array_equal(a["offsets"], 8 * np.arange(1, s.n + 1))
    assert np.array_equal(a["types"], np.full(s.n, 11))       # VTK_VOXEL
    assert np.array_equal(a["processor"], np.zeros(s.n))
    # branches: the root's children in key order
    bnodes = np.nonzero(t.level == 1)[0]
    bnodes = bnodes[np.argsort(t.key[bnodes], kind="stable")]
    piece, a = _vtu(tmp_path / "vtk" / "branches_000003.000000.vtu")
    assert piece.get("NumberOfCells") == str(len(bnodes)) == "8"
    corners, desc, ch, ach = _expected_boxes(orc, d, t, bnodes)
    assert np.array_equal(a["corners"], corners)
    assert np.array_equal(a["level"], np.ones(len(bnodes)))
    assert np.array_equal(a["leaves"], t.leaves[bnodes])
    assert np.array_equal(a["descendants"], desc)
    assert desc.sum() + len(bnodes) + 1 == t.nnodes          # every node but the root
    assert np.array_equal(a["charge"], ch) and np.array_equal(a["abs_charge"], ach)
    # after pepc_timber_tree: no tree, no file
    assert "vtk_write_leaves(): tree is not allocated" in r.stdout
    assert not (tmp_path / "vtk" / "leaves_000004.000000.vtu").exists()
