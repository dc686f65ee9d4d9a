"""GPU parity on tiny systems (a handful of particles up to 1000): one group
or a few, fewer groups than walk work queues, a root that is a twig with
leaf children only, groups straddling the 64-lane slicing.  Counts and walk
tiles bit-exact against the oracle, FP64 forces to 1e-12, FP32 to rounding
level.  A single particle has a zero-extent box, which the library reports
(PEPC_GPU_EBOX) instead of building a tree."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

THETA2, EPS2 = 0.36, 1e-4


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pepc_amd.solver import Solver
    return Solver


def _system(n):
    rng = np.random.default_rng(1000 + n)
    x, y, z = rng.random((3, n))
    q = rng.uniform(-1.0, 1.0, n)
    return x, y, z, q


@pytest.mark.parametrize("G,pack", [(1, 0), (16, 32), (64, 64)])
@pytest.mark.parametrize("n", [2, 3, 5, 9, 17, 64, 65, 100, 1000])
def test_tiny_system_matches_oracle(gpu, orc, n, G, pack):
    x, y, z, q = _system(n)
    d = orc.domain_decompose(x, y, z, q)
    t = orc.tree_build(d.keys, d.x, d.y, d.z, d.q, d.box)
    w = orc.walk_group(d.x, d.y, d.z, t, THETA2, EPS2, tile_size=G, keys=d.keys, pack=pack)
    print(f"[tiny n={n} G={G} pack={pack}] groups={len(w.tile_first)}", flush=True)
    for prec in ("fp64", "fp32"):
        sol = gpu(theta2=THETA2, eps2=EPS2, group_size=G, group_pack=pack, precision=prec)
        st = sol.grow(x, y, z, q)
        res = sol.traverse()
        wk = sol.read_walk()
        assert st["ngroups"] == len(w.tile_first)
        assert np.array_equal(wk["gfirst"].astype(np.int64), w.tile_first)
        assert np.array_equal(wk["gcount"].astype(np.int64), w.tile_count)
        assert np.array_equal(wk["p2p"].astype(np.int32), w.p2p)
        assert np.array_equal(wk["m2p"].astype(np.int32), w.m2p)
        assert wk["stats"]["interactions"] == w.num_interactions
        assert wk["stats"]["mac_evaluations"] == w.num_mac
        e = res.e.cpu().numpy().T
        pot = res.pot.cpu().numpy()
        tol = 1e-12 if prec == "fp64" else 1e-5
        assert np.max(np.abs(e - w.e)) <= tol * np.abs(w.e).max()
        assert np.max(np.abs(pot - w.pot)) <= tol * np.abs(w.pot).max()
        sol.close()


def test_single_particle_is_reported(gpu):
    from pepc_amd import _lib
    sol = gpu(theta2=THETA2, eps2=EPS2)
    with pytest.raises(_lib.PepcGpuError) as ei:
        sol.grow(*(np.array([v]) for v in (0.5, 0.5, 0.5, 1.0)))
    assert ei.value.code == _lib.EBOX
    sol.close()
