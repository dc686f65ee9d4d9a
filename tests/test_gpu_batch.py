"""Batched systems: one grow/traverse over many independent systems equals a
separate solve of each system, bit for bit (keys, permutation, positions after
the duplicate fix-up, node counts, walk groups, interaction and MAC counts,
fields and potentials), and every system equals the oracle's solve of that
system alone (integer work bit-exact, FP64 fields to 1e-12)."""
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


def _batch_systems():
    systems = workloads.ensemble(nsys=10, seed=41, nmin=64, nmax=4096)
    r = np.random.default_rng(42)
    # two particles
    p = r.random((2, 3))
    systems.append(workloads.System("pair", p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(),
                                    np.array([1.0, -1.0]), np.ones(2), np.zeros((2, 3)),
                                    eps2=1e-4, theta2=0.36))
    # an empty system between two others
    systems.insert(4, workloads.System("empty", *(np.zeros(0) for _ in range(5)),
                                       np.zeros((0, 3)), eps2=1e-4, theta2=0.36))
    # coincident pairs: exercises the duplicate-key fix-up inside one range
    p = r.random((700, 3))
    p[300:306] = p[0:6]
    systems.insert(7, workloads.System("coincident", p[:, 0].copy(), p[:, 1].copy(),
                                       p[:, 2].copy(), np.ones(700), np.ones(700),
                                       np.zeros((700, 3)), eps2=1e-4, theta2=0.36))
    return systems


@pytest.mark.parametrize("precision,G,pack", [("fp32", 64, 64), ("fp64", 16, 0)])
def test_batch_equals_separate_solves(gpu, precision, G, pack):
    systems = _batch_systems()
    x, y, z, q, m, v, off = workloads.concat(systems)
    print(f"[batch {precision} G={G}] nsys={len(systems)} n={off[-1]}", flush=True)
    sol = gpu(theta2=0.36, eps2=1e-4, group_size=G, precision=precision, group_pack=pack)
    bst = sol.grow_batch(x, y, z, q, off)
    bres = sol.traverse()
    bkeys, bperm, bx, by, bz, bq = sol.sorted_host()
    bwalk = sol.read_walk()
    be = bres.e.cpu().numpy()
    bpot = bres.pot.cpu().numpy()
    bwork = bres.work.cpu().numpy()
    bst = bwalk["stats"]

    one = gpu(theta2=0.36, eps2=1e-4, group_size=G, precision=precision, group_pack=pack)
    tot = {"nnodes": 0, "ntwig": 0, "ngroups": 0, "interactions": 0, "mac_evaluations": 0,
           "dup_fixed": 0}
    gfirst = bwalk["gfirst"].astype(np.int64)
    for k, s in enumerate(systems):
        a, b = off[k], off[k + 1]
        if b == a:
            assert not np.any((gfirst >= a) & (gfirst < b))
            continue
        st = one.grow(s.x, s.y, s.z, s.q)
        res = one.traverse()
        keys, perm, xs, ys, zs, qs = one.sorted_host()
        w = one.read_walk()
        assert np.array_equal(bkeys[a:b], keys), s.name
        assert np.array_equal(bperm[a:b] - a, perm), s.name
        assert np.array_equal(bx[a:b], xs) and np.array_equal(by[a:b], ys)
        assert np.array_equal(bz[a:b], zs) and np.array_equal(bq[a:b], qs)
        sel = (gfirst >= a) & (gfirst < b)
        assert np.array_equal(gfirst[sel] - a, w["gfirst"].astype(np.int64)), s.name
        assert np.array_equal(bwalk["gcount"][sel], w["gcount"])
        assert np.array_equal(bwalk["gmac"][sel], w["gmac"])
        assert np.array_equal(bwalk["p2p"][a:b], w["p2p"])
        assert np.array_equal(bwalk["m2p"][a:b], w["m2p"])
        assert np.array_equal(be[:, a:b], res.e.cpu().numpy()), s.name
        assert np.array_equal(bpot[a:b], res.pot.cpu().numpy())
        assert np.array_equal(bwork[a:b], res.work.cpu().numpy())
        ws = w["stats"]
        for key in tot:
            tot[key] += ws[key] if key in ws else st[key]
    for key, val in tot.items():
        assert bst[key] == val, key
    sol.close()
    one.close()


@pytest.mark.parametrize("G,pack", [(16, 0), (32, 32), (64, 64)])
def test_batch_matches_oracle_per_system(gpu, orc, G, pack):
    systems = _batch_systems()
    x, y, z, q, m, v, off = workloads.concat(systems)
    sol = gpu(theta2=0.36, eps2=1e-4, group_size=G, precision="fp64", group_pack=pack)
    sol.grow_batch(x, y, z, q, off)
    res = sol.traverse()
    keys, perm, xs, ys, zs, qs = sol.sorted_host()
    wk = sol.read_walk()
    gf = wk["gfirst"].astype(np.int64)
    e, pot, work = res.e.cpu().numpy().T, res.pot.cpu().numpy(), res.work.cpu().numpy()
    for k, s in enumerate(systems):
        a, b = off[k], off[k + 1]
        if b - a < 2:
            continue
        d = orc.domain_decompose(s.x, s.y, s.z, s.q)
        t = orc.tree_build(d.keys, d.x, d.y, d.z, d.q, d.box)
        w = orc.walk_group(d.x, d.y, d.z, t, 0.36, 1e-4, tile_size=G, keys=d.keys, pack=pack)
        assert np.array_equal(keys[a:b], d.keys), s.name
        assert np.array_equal(perm[a:b] - a, d.perm), s.name
        assert np.array_equal(xs[a:b], d.x) and np.array_equal(zs[a:b], d.z), s.name
        sel = (gf >= a) & (gf < b)
        assert np.array_equal(gf[sel] - a, w.tile_first), s.name
        assert np.array_equal(wk["gcount"][sel].astype(np.int64), w.tile_count), s.name
        assert np.array_equal(wk["gmac"][sel].astype(np.int64), w.tile_mac), s.name
        assert np.array_equal(wk["p2p"][a:b].astype(np.int32), w.p2p), s.name
        assert np.array_equal(wk["m2p"][a:b].astype(np.int32), w.m2p), s.name
        assert np.array_equal(work[a:b], w.work), s.name
        assert np.max(np.abs(e[a:b] - w.e)) <= 1e-12 * np.abs(w.e).max(), s.name
        assert np.max(np.abs(pot[a:b] - w.pot)) <= 1e-12 * np.abs(w.pot).max(), s.name
    sol.close()


def test_batch_rejects_bad_offsets_and_degenerate_system(gpu):
    from pepc_amd import _lib
    systems = workloads.ensemble(nsys=3, seed=43, nmin=64, nmax=256)
    x, y, z, q, m, v, off = workloads.concat(systems)
    sol = gpu(theta2=0.36, eps2=1e-4)
    bad = off.copy()
    bad[1], bad[2] = bad[2], bad[1]
    with pytest.raises(_lib.PepcGpuError) as ei:
        sol.grow_batch(x, y, z, q, bad)
    assert ei.value.code == _lib.EINVAL
    bad = off.copy()
    bad[-1] -= 1
    with pytest.raises(ValueError):      # offsets must cover the arrays
        sol.grow_batch(x, y, z, q, bad)
    # a one-particle system has a zero-size box (module_box.f90:80-84 would
    # divide by zero): reported as a degenerate box
    one = np.array([0, 1, off[-1] + 1], np.int64)
    xx = np.concatenate([[0.5], x])
    with pytest.raises(_lib.PepcGpuError) as ei:
        sol.grow_batch(xx, np.concatenate([[0.5], y]), np.concatenate([[0.5], z]),
                       np.concatenate([[1.0], q]), one)
    assert ei.value.code == _lib.EBOX
    # the context stays usable
    st = sol.grow_batch(x, y, z, q, off)
    assert st["nleaf"] == off[-1]
    sol.close()
