"""Oracle parity at the benchmark sizes.  The oracle runs the whole solve at
these sizes in seconds (its walk is threaded), so every target is compared:

* C2 Coulomb explosion, N = 4,194,304, theta = 0.6, at the benchmark's walk
  rule (G = 16, sibling tiles packed to 32): keys, perm, node and twig
  counts, the groups, per-group MAC counts, per-target P2P/M2P counts and
  work bit-exact; FP64 fields and potentials within 1e-12 of the maximum;
  the FP32 + Kahan path's counts bit-exact too (same acceptance decisions);
* C3 clustered gravity, N = 8,388,608, attractive sign, theta = 0.5, at the
  reference's tile rule (G = 16): the same comparisons;
* C4 256 batched systems at the rule the autotuner installs: every system's
  keys, perm, groups and counts bit-exact against a separate oracle solve,
  FP64 fields within 1e-12.

The reference rule (G = 16, no packing) at small sizes is in
test_gpu_parity.py; these runs pin the rules and sizes the benchmark uses."""
import time

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


def _check_against_oracle(orc, sol, s, G, pack, label, fields=True):
    t0 = time.time()
    d = orc.domain_decompose(s.x, s.y, s.z, s.q)
    t = orc.tree_build(d.keys, d.x, d.y, d.z, d.q, d.box)
    w = orc.walk_group(d.x, d.y, d.z, t, s.theta2, s.eps2, tile_size=G, sign=s.sign,
                       keys=d.keys, pack=pack)
    print(f"[{label}] oracle solve {time.time() - t0:.1f} s: {t.nnodes} nodes, "
          f"{w.num_interactions} interactions", flush=True)
    st = sol.sync()
    keys, perm, xs, ys, zs, qs = sol.sorted_host()
    assert np.array_equal(keys, d.keys), label
    assert np.array_equal(perm, d.perm), label
    assert np.array_equal(xs, d.x) and np.array_equal(ys, d.y) and np.array_equal(zs, d.z)
    assert st["nnodes"] == t.nnodes and st["ntwig"] == t.ntwig, label
    wk = sol.read_walk()
    assert np.array_equal(wk["gfirst"].astype(np.int64), w.tile_first), label
    assert np.array_equal(wk["gcount"].astype(np.int64), w.tile_count), label
    assert np.array_equal(wk["gmac"].astype(np.int64), w.tile_mac), label
    assert np.array_equal(wk["p2p"].astype(np.int32), w.p2p), label
    assert np.array_equal(wk["m2p"].astype(np.int32), w.m2p), label
    assert wk["stats"]["interactions"] == w.num_interactions
    assert wk["stats"]["mac_evaluations"] == w.num_mac
    return w


def _traverse_fp64_and_compare(orc, gpu, s, G, pack, label, X):
    sol = gpu(theta2=s.theta2, eps2=s.eps2, sign=s.sign, group_size=G, group_pack=pack,
              precision="fp64")
    sol.grow(*X)
    res = sol.traverse()
    w = _check_against_oracle(orc, sol, s, G, pack, label)
    assert np.array_equal(res.work.cpu().numpy(), w.work)
    e = res.e.cpu().numpy().T
    pot = res.pot.cpu().numpy()
    de = np.max(np.abs(e - w.e)) / np.abs(w.e).max()
    dp = np.max(np.abs(pot - w.pot)) / np.abs(w.pot).max()
    print(f"[{label}] FP64 fields vs oracle over all {s.n} targets: max {de:.1e} (field), "
          f"{dp:.1e} (potential) of the maximum", flush=True)
    assert de <= 1e-12 and dp <= 1e-12
    sol.close()
    # FP32 + Kahan: the same decisions, so the same counts
    s32 = gpu(theta2=s.theta2, eps2=s.eps2, sign=s.sign, group_size=G, group_pack=pack,
              precision="fp32")
    s32.grow(*X)
    s32.traverse()
    wk = s32.read_walk()
    assert np.array_equal(wk["p2p"].astype(np.int32), w.p2p)
    assert np.array_equal(wk["m2p"].astype(np.int32), w.m2p)
    assert np.array_equal(wk["gmac"].astype(np.int64), w.tile_mac)
    s32.close()


def test_c2_explosion_oracle_parity_at_bench_rule(gpu, orc):
    import torch
    s = workloads.explosion()
    X = [torch.from_numpy(a).to("cuda") for a in (s.x, s.y, s.z, s.q)]
    _traverse_fp64_and_compare(orc, gpu, s, 16, 32, "C2 4M theta=0.6 G=16 pack=32", X)


def test_c3_plummer_oracle_parity(gpu, orc):
    import torch
    s = workloads.plummer()
    X = [torch.from_numpy(a).to("cuda") for a in (s.x, s.y, s.z, s.q)]
    _traverse_fp64_and_compare(orc, gpu, s, 16, 0, "C3 8M theta=0.5 G=16 sign=-1", X)


def test_c4_ensemble_oracle_parity_at_autotuned_rule(gpu, orc):
    systems = workloads.ensemble()
    x, y, z, q, m, v, off = workloads.concat(systems)
    tune = gpu(theta2=0.36, eps2=1e-4, precision="fp32")
    best, _ = tune.autotune(x, y, z, q, offsets=off, reps=1)
    cfg = tune.config()
    G, pack = cfg["group_size"], cfg["group_pack"]
    print(f"[C4] autotuned rule G={G} pack={pack} list_tile={cfg['list_tile']}", flush=True)
    # the tuned FP32 context's own counts, and an FP64 context at its rule
    tune.grow_batch(x, y, z, q, off)
    tune.traverse()
    w32 = tune.read_walk()
    sol = gpu(theta2=0.36, eps2=1e-4, precision="fp64", group_size=G, group_pack=pack)
    sol.grow_batch(x, y, z, q, off)
    res = sol.traverse()
    keys, perm, xs, ys, zs, qs = sol.sorted_host()
    wk = sol.read_walk()
    gf = wk["gfirst"].astype(np.int64)
    e, pot, work = res.e.cpu().numpy().T, res.pot.cpu().numpy(), res.work.cpu().numpy()
    t0 = time.time()
    worst = 0.0
    for k, s in enumerate(systems):
        a, b = off[k], off[k + 1]
        d = orc.domain_decompose(s.x, s.y, s.z, s.q)
        t = orc.tree_build(d.keys, d.x, d.y, d.z, d.q, d.box)
        w = orc.walk_group(d.x, d.y, d.z, t, s.theta2, s.eps2, tile_size=G, keys=d.keys, pack=pack)
        assert np.array_equal(keys[a:b], d.keys), k
        assert np.array_equal(perm[a:b] - a, d.perm), k
        sel = (gf >= a) & (gf < b)
        assert np.array_equal(gf[sel] - a, w.tile_first), k
        assert np.array_equal(wk["gcount"][sel].astype(np.int64), w.tile_count), k
        assert np.array_equal(wk["gmac"][sel].astype(np.int64), w.tile_mac), k
        assert np.array_equal(w32["gmac"][sel].astype(np.int64), w.tile_mac), k
        for src in (wk, w32):
            assert np.array_equal(src["p2p"][a:b].astype(np.int32), w.p2p), k
            assert np.array_equal(src["m2p"][a:b].astype(np.int32), w.m2p), k
        assert np.array_equal(work[a:b], w.work), k
        de = np.max(np.abs(e[a:b] - w.e)) / np.abs(w.e).max()
        dp = np.max(np.abs(pot[a:b] - w.pot)) / np.abs(w.pot).max()
        worst = max(worst, de, dp)
    print(f"[C4] {len(systems)} systems vs oracle in {time.time() - t0:.1f} s; worst FP64 "
          f"difference {worst:.1e} of the system's maximum", flush=True)
    assert worst <= 1e-12
    tune.close()
    sol.close()
