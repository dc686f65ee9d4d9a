"""Particle checkpoints of the Fortran shim in the reference's formats
(src/utils/module_checkpoint.f90).  The expected bytes are built here from
the reference's type map: the MPI-IO file is a 128-byte header {int64
n_total, int32 itime} followed by one 96-byte record per particle
{x(3), work, key, node_leaf, label, q, e(3), pot} (module_pepc_types.f90:49-58,
107-147; all components 8 bytes, module_pepc_kinds.f90:63-70); the binary
file holds the records alone.  Runs on the CPU (no GPU context is created).
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "pepc_amd", "fortran", "pepc_mini_gpu")

REC = np.dtype([("x", "<f8", (3,)), ("work", "<f8"), ("key", "<i8"), ("node_leaf", "<i8"),
                ("label", "<i8"), ("q", "<f8"), ("e", "<f8", (3,)), ("pot", "<f8")])


def test_record_layout_is_96_bytes():
    assert REC.itemsize == 96


def test_checkpoint_formats_roundtrip(tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("Fortran frontend not built (amdflang absent)")
    n = 1000
    r = np.random.default_rng(61)
    x, y, z = r.random((3, n))
    q = r.normal(size=n)
    fin = tmp_path / "in.bin"
    with open(fin, "wb") as f:
        np.array([n], np.int64).tofile(f)
        for a in (x, y, z, q):
            a.tofile(f)
    out = tmp_path / "out.bin"
    res = subprocess.run([EXE, "--checkpoint", str(fin), str(out)], cwd=tmp_path,
                         capture_output=True, text=True, timeout=60)
    assert res.returncode == 0, res.stdout + res.stderr
    i = np.arange(1, n + 1)
    want = np.zeros(n, REC)
    want["x"] = np.stack([x, y, z], 1)
    want["work"] = 0.5 * i
    want["key"] = 1000 + i
    want["node_leaf"] = i
    want["label"] = i - 1
    want["q"] = q
    want["e"] = np.stack([i, 2.0 * i, 3.0 * i], 1)
    want["pot"] = -i
    mpi = (tmp_path / "particles" / "mpi" / "particle_000000000007.mpi").read_bytes()
    header = np.array([n], "<i8").tobytes() + np.array([7], "<i4").tobytes()
    assert mpi[:12] == header and mpi[12:128] == bytes(116)
    assert mpi[128:] == want.tobytes()
    binf = (tmp_path / "particles" / "binary" / "particle_000000000007_000000.bin").read_bytes()
    assert binf == want.tobytes()
    assert (tmp_path / "particles" / "mpi" / "particle_000000000007.mpi.h").exists()
    back = out.read_bytes()
    assert np.frombuffer(back[:16], "<i8").tolist() == [n, 7]
    assert back[16:] == want.tobytes()
