"""Autotuner and the tuning parameters it chooses between.

* the persistent grid size never changes a result (bit-identical);
* the list tile changes only the FP summation batches: identical interaction
  counts, forces equal to rounding;
* the group size and pack change the interaction lists (the declared group
  rule), and each choice still matches the oracle at that rule
  (test_gpu_parity.py);
* autotune returns the fastest candidate, installs it, and a solve with the
  installed configuration equals a solve with that configuration given
  explicitly.
"""
import numpy as np
import pytest

from pepc_amd import workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pepc_amd.solver import Solver
    return Solver


def _solve(sol, s):
    sol.grow(s.x, s.y, s.z, s.q)
    res = sol.traverse()
    w = sol.read_walk()
    return res.e.cpu().numpy(), res.pot.cpu().numpy(), w


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_grid_and_list_tile(gpu, precision):
    s = workloads.explosion(n=65536, seed=71)
    kw = dict(theta2=s.theta2, eps2=s.eps2, precision=precision, group_size=64, group_pack=64)
    e0, p0, w0 = _solve(gpu(**kw), s)
    e1, p1, w1 = _solve(gpu(walk_blocks=37, **kw), s)
    assert np.array_equal(e0, e1) and np.array_equal(p0, p1)
    other = 64 if precision == "fp32" else 32
    e2, p2, w2 = _solve(gpu(list_tile=other, **kw), s)
    assert np.array_equal(w0["p2p"], w2["p2p"]) and np.array_equal(w0["m2p"], w2["m2p"])
    rel = np.linalg.norm(e2 - e0, axis=0) / np.linalg.norm(e0, axis=0)
    print(f"[{precision}] list tile {other}: median rel change {np.median(rel):.2e} "
          f"max {rel.max():.2e}", flush=True)
    if precision == "fp32":     # FP32 rounding of differently batched sums
        assert np.median(rel) < 1e-6 and rel.max() < 1e-4
    else:
        assert rel.max() < 1e-12


def test_bad_list_tile_rejected(gpu):
    from pepc_amd import _lib
    with pytest.raises(_lib.PepcGpuError):
        gpu(precision="fp32", list_tile=32)
    with pytest.raises(_lib.PepcGpuError):
        gpu(precision="fp64", list_tile=128)


def test_autotune_installs_fastest(gpu):
    s = workloads.explosion(n=131072, seed=72)
    sol = gpu(theta2=s.theta2, eps2=s.eps2, precision="fp32")
    best, cands = sol.autotune(s.x, s.y, s.z, s.q, reps=1)
    print(f"best {best}", flush=True)
    assert len(cands) == 36
    tot = [c["t_grow_ms"] + c["t_walk_ms"] for c in cands]
    assert all(t > 0 for t in tot)
    assert best["t_grow_ms"] + best["t_walk_ms"] == pytest.approx(min(tot))
    cfg = sol.config()
    for k in ("group_size", "group_pack", "list_tile", "walk_blocks"):
        assert cfg[k] == best[k]
    e0, p0, w0 = _solve(sol, s)
    ref = gpu(theta2=s.theta2, eps2=s.eps2, precision="fp32", group_size=best["group_size"],
              group_pack=best["group_pack"], list_tile=best["list_tile"],
              walk_blocks=best["walk_blocks"])
    e1, p1, w1 = _solve(ref, s)
    assert np.array_equal(e0, e1) and np.array_equal(p0, p1)


def test_autotune_batch(gpu):
    systems = workloads.ensemble(nsys=16, seed=73, nmin=1024, nmax=8192)
    x, y, z, q, m, v, off = workloads.concat(systems)
    sol = gpu(theta2=0.36, eps2=1e-4, precision="fp32")
    best, cands = sol.autotune(x, y, z, q, offsets=off, reps=1)
    print(f"batch best {best}", flush=True)
    st = sol.grow_batch(x, y, z, q, off)
    assert st["nleaf"] == off[-1]
    sol.traverse()
