"""Equal keys on the GPU: the permutation, the duplicate-key fix-up and
everything downstream must equal the oracle bit for bit when particles share
positions.

The oracle's presort (or_sort_radix_ip) is pinned against the reference's own
sort_radix_ip, compiled from sl_pepc's sources, including the order it gives
to equal keys (tests/test_oracle_sort.py).  That order decides which particle
of a tied run keeps its key and which one the fix-up (module_domains.f90:
248-307) moves, so with mixed +-1 charges the fields depend on it.  Compared
here through the C-ABI: sorted keys after the fix-up, perm, sorted positions
in sorted order (not as a set), charges, tree size, per-target P2P/M2P
counts, work (all bit-exact) and FP64 fields (1e-12 of the maximum)."""
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


CASES = {
    # pairs and triples all over the unit cube: ties in buckets of every size
    "pairs20k": dict(n=20000, seed=8),
    # a 2e-4 cube holding a third of the particles: the buckets that hold the
    # ties stay above 256 elements for several digits, so the replay recurses
    "cluster12k": dict(n=12000, seed=9, cluster=4000),
    # a large top-level pass (one rotor walk over 262,144 elements)
    "sparse262k": dict(n=262144, seed=10, pairs=0.004, triples=0.0005),
}


def _compare(orc, sol, s, d, label):
    keys, perm, x, y, z, q = sol.sorted_host()
    assert np.array_equal(keys, d.keys), label
    assert np.array_equal(perm, d.perm), label
    assert np.array_equal(x, d.x) and np.array_equal(y, d.y) and np.array_equal(z, d.z), label
    assert np.array_equal(q, d.q), label


@pytest.mark.parametrize("name", list(CASES))
def test_tied_keys_bitexact(gpu, orc, name):
    s = workloads.coincident(**CASES[name])
    d = orc.domain_decompose(s.x, s.y, s.z, s.q)
    t = orc.tree_build(d.keys, d.x, d.y, d.z, d.q, d.box)
    w = orc.walk_group(d.x, d.y, d.z, t, s.theta2, s.eps2, tile_size=16, keys=d.keys)
    print(f"[{name}] n={s.n} moved by the fix-up: {d.nfixed}", flush=True)
    assert d.nfixed > 0
    sol = gpu(theta2=s.theta2, eps2=s.eps2, group_size=16, precision="fp64")
    st = sol.grow(s.x, s.y, s.z, s.q)
    res = sol.traverse()
    tg, _ = sol.timings()
    print(f"[{name}] grow {tg:.2f} ms", flush=True)
    assert st["dup_fixed"] == d.nfixed
    _compare(orc, sol, s, d, name)
    assert st["nnodes"] == t.nnodes and st["ntwig"] == t.ntwig
    wk = sol.read_walk()
    assert np.array_equal(wk["p2p"].astype(np.int32), w.p2p)
    assert np.array_equal(wk["m2p"].astype(np.int32), w.m2p)
    assert np.array_equal(res.work.cpu().numpy(), w.work)
    e = res.e.cpu().numpy().T
    assert np.max(np.abs(e - w.e)) <= 1e-12 * np.abs(w.e).max()
    assert np.max(np.abs(res.pot.cpu().numpy() - w.pot)) <= 1e-12 * np.abs(w.pot).max()
    # the same context again (buffers reused, fresh replay)
    sol.grow(s.x, s.y, s.z, s.q)
    _compare(orc, sol, s, d, name + " (second grow)")
    sol.close()


def test_tied_keys_in_batched_systems(gpu, orc):
    # each system's ties follow its own presort (one sort per system), as
    # separate reference runs would order them
    systems = [workloads.coincident(n=n, seed=60 + k, cluster=c)
               for k, (n, c) in enumerate([(3000, 0), (700, 300), (50, 0), (9000, 3000), (257, 0)])]
    systems.insert(2, workloads.mini(n=2000, seed=66))   # a system without ties
    x, y, z, q, m, v, off = workloads.concat(systems)
    sol = gpu(theta2=0.36, eps2=1e-4, group_size=16, precision="fp64")
    st = sol.grow_batch(x, y, z, q, off)
    keys, perm, xs, ys, zs, qs = sol.sorted_host()
    fixed = 0
    for k, s in enumerate(systems):
        a, b = off[k], off[k + 1]
        d = orc.domain_decompose(s.x, s.y, s.z, s.q)
        fixed += d.nfixed
        assert np.array_equal(keys[a:b], d.keys), k
        assert np.array_equal(perm[a:b] - a, d.perm), k
        assert np.array_equal(xs[a:b], d.x) and np.array_equal(ys[a:b], d.y), k
        assert np.array_equal(zs[a:b], d.z) and np.array_equal(qs[a:b], d.q), k
    assert st["dup_fixed"] == fixed > 0
    sol.close()


def test_equal_keys_distinct_positions(gpu, orc):
    # particles in one 2^-20 cell at different positions: equal keys that the
    # fix-up separates, positions of the moved ones rewritten to cell centres
    r = np.random.default_rng(12)
    p = r.random((5000, 3))
    cell = 1.0 / 2 ** 20
    p[4000:4040] = p[0:40] + 0.3 * cell * r.random((40, 3))
    q = r.choice([-1.0, 1.0], 5000)
    d = orc.domain_decompose(p[:, 0], p[:, 1], p[:, 2], q)
    sol = gpu(theta2=0.36, eps2=1e-4, group_size=16, precision="fp64")
    st = sol.grow(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), q)
    assert st["dup_fixed"] == d.nfixed > 0
    _compare(orc, sol, None, d, "one-cell runs")
    sol.close()


@pytest.mark.parametrize("runs", [(257,), (300, 700), (1500, 260, 2)])
def test_long_runs_of_equal_keys(gpu, orc, runs):
    """Runs of identical positions longer than the reference sort's
    insertion threshold (256, sl_tune_intern.h:30): rs_rec_ip recurses
    through every digit with the whole run in one bucket, and the fix-up then
    spreads the run over consecutive keys.  Keys, perm, fixed-up positions,
    tree size, per-target counts, work bit-exact; FP64 fields to 1e-12."""
    r = np.random.default_rng(70 + len(runs))
    n = 4000 + sum(runs)
    p = r.random((n, 3))
    at = r.permutation(n)
    k = 0
    for length in runs:
        src = at[k]
        p[at[k + 1:k + length]] = p[src]
        k += length
    q = r.choice(np.array([-1.0, 1.0]), n)
    x, y, z = (p[:, j].copy() for j in range(3))
    d = orc.domain_decompose(x, y, z, q)
    t = orc.tree_build(d.keys, d.x, d.y, d.z, d.q, d.box)
    w = orc.walk_group(d.x, d.y, d.z, t, 0.36, 1e-4, tile_size=16, keys=d.keys)
    print(f"[long runs {runs}] n={n} moved by the fix-up: {d.nfixed}", flush=True)
    assert d.nfixed >= sum(runs) - 2 * len(runs)
    sol = gpu(theta2=0.36, eps2=1e-4, group_size=16, precision="fp64", node_factor=8.0)
    st = sol.grow(x, y, z, q)
    res = sol.traverse()
    assert st["dup_fixed"] == d.nfixed
    _compare(orc, sol, None, d, f"runs {runs}")
    assert st["nnodes"] == t.nnodes and st["ntwig"] == t.ntwig
    wk = sol.read_walk()
    assert np.array_equal(wk["p2p"].astype(np.int32), w.p2p)
    assert np.array_equal(wk["m2p"].astype(np.int32), w.m2p)
    assert np.array_equal(res.work.cpu().numpy(), w.work)
    e = res.e.cpu().numpy().T
    assert np.max(np.abs(e - w.e)) <= 1e-12 * np.abs(w.e).max()
    assert np.max(np.abs(res.pot.cpu().numpy() - w.pot)) <= 1e-12 * np.abs(w.pot).max()
    sol.close()
