"""Randomized batched sweep: seeded batches of 1 to 40 independent systems
(empty systems, two-particle systems, sizes up to 5,000; uniform, a dense
cluster, coincident pairs; each in its own box of random size and place),
with random opening angle, walk group rule, sign and softening, grown and
traversed in one call.  Every system equals the oracle's solve of that system
alone: keys, permutation, fixed-up positions, groups, per-target P2P/M2P
counts, MAC counts and work bit-exact; FP64 fields and potentials to 1e-12 of
that system's maximum.  The batch's node, group and interaction totals are
the sums of the oracle's per-system ones."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = 12


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pepc_amd.solver import Solver
    return Solver


def _system(rng):
    u = rng.random()
    if u < 0.08:
        n = 0
    elif u < 0.16:
        n = 2
    else:
        n = int(np.exp(rng.uniform(np.log(3.0), np.log(5000.0))))
    kind = int(rng.integers(0, 3))
    p = rng.random((n, 3))
    if kind == 1 and n >= 6:
        k = n // 3
        p[:k] = 0.4 + 1e-3 * rng.random((k, 3))
    if kind == 2 and n >= 20:
        m = n // 10
        idx = rng.permutation(n)
        p[idx[m:2 * m]] = p[idx[:m]]
    p = rng.uniform(-10.0, 10.0, 3) + p * rng.uniform(0.01, 5.0)
    q = rng.choice(np.array([-1.0, 1.0]), n) if rng.random() < 0.5 else rng.normal(size=n)
    return p, q


def _case(seed):
    rng = np.random.default_rng(2000 + seed)
    nsys = int(rng.integers(1, 41))
    systems = [_system(rng) for _ in range(nsys)]
    if sum(len(q) for _, q in systems) == 0:      # the batch itself must not be empty
        systems.append((rng.random((2, 3)), np.array([1.0, -1.0])))
    theta2 = float(rng.choice([0.0, 0.09, 0.36, 0.81]))
    G = int(rng.choice([1, 8, 16, 32, 64]))
    pack = int(rng.choice([0, min(64, 2 * G)])) if 1 < G < 64 else 0
    sign = float(rng.choice([1.0, -1.0]))
    eps2 = float(rng.choice([0.0, 1e-4]))
    return systems, dict(theta2=theta2, G=G, pack=pack, sign=sign, eps2=eps2)


@pytest.mark.parametrize("seed", range(CASES))
def test_random_batch_parity(gpu, orc, seed):
    systems, c = _case(seed)
    sizes = [len(q) for _, q in systems]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    p = np.concatenate([s[0] for s in systems]) if off[-1] else np.zeros((0, 3))
    q = np.concatenate([s[1] for s in systems])
    x, y, z = (np.ascontiguousarray(p[:, k]) for k in range(3))
    print(f"[batch seed {seed}] nsys={len(systems)} n={off[-1]} sizes={sizes} {c}", flush=True)
    sol = gpu(theta2=c["theta2"], eps2=c["eps2"], group_size=c["G"], precision="fp64",
              sign=c["sign"], group_pack=c["pack"])
    st = sol.grow_batch(x, y, z, q, off)
    res = sol.traverse()
    keys, perm, xs, ys, zs, qs = sol.sorted_host()
    wk = sol.read_walk()
    gf = wk["gfirst"].astype(np.int64)
    e, pot, work = res.e.cpu().numpy().T, res.pot.cpu().numpy(), res.work.cpu().numpy()
    nnodes = ntwig = ngroups = inter = 0
    for k, (ps, qk) in enumerate(systems):
        a, b = off[k], off[k + 1]
        sel = (gf >= a) & (gf < b)
        if b == a:
            assert not sel.any()
            continue
        d = orc.domain_decompose(ps[:, 0].copy(), ps[:, 1].copy(), ps[:, 2].copy(), qk)
        t = orc.tree_build(d.keys, d.x, d.y, d.z, d.q, d.box)
        w = orc.walk_group(d.x, d.y, d.z, t, c["theta2"], c["eps2"], tile_size=c["G"],
                           sign=c["sign"], keys=d.keys, pack=c["pack"])
        tag = f"system {k} n={b - a}"
        assert np.array_equal(keys[a:b], d.keys), tag
        assert np.array_equal(perm[a:b] - a, d.perm), tag
        assert np.array_equal(xs[a:b], d.x) and np.array_equal(ys[a:b], d.y), tag
        assert np.array_equal(zs[a:b], d.z) and np.array_equal(qs[a:b], d.q), tag
        assert np.array_equal(gf[sel] - a, w.tile_first), tag
        assert np.array_equal(wk["gcount"][sel].astype(np.int64), w.tile_count), tag
        assert np.array_equal(wk["gmac"][sel].astype(np.int64), w.tile_mac), tag
        assert np.array_equal(wk["p2p"][a:b].astype(np.int32), w.p2p), tag
        assert np.array_equal(wk["m2p"][a:b].astype(np.int32), w.m2p), tag
        assert np.array_equal(work[a:b], w.work), tag
        assert np.max(np.abs(e[a:b] - w.e)) <= 1e-12 * max(np.abs(w.e).max(), 1e-300), tag
        assert np.max(np.abs(pot[a:b] - w.pot)) <= 1e-12 * max(np.abs(w.pot).max(), 1e-300), tag
        nnodes += t.nnodes
        ntwig += t.ntwig
        ngroups += len(w.tile_first)
        inter += int(w.p2p.sum() + w.m2p.sum())
    stats = wk["stats"]
    assert st["nnodes"] == nnodes and st["ntwig"] == ntwig
    assert stats["ngroups"] == ngroups and stats["interactions"] == inter
    sol.close()
