"""Randomized sink sweep: seeded source systems (uniform, a dense cluster,
coincident pairs) with random opening angle, softening and sign, and sink
sets that mix points inside and outside the source box, points exactly on
source particles (the reference's zero-separation self rule,
module_tree_walk_simple.f90:331-340), points a few ulps off a source,
repeated sinks and sinks in a cluster of their own.  Per-sink mode, one sink
per tile (:155-159), against the oracle: per-sink P2P/M2P counts and work
bit-exact, FP64 fields and potentials to 1e-12 of each sink's sum of
absolute terms (sum_j |q_j| / r^2 and sum_j |q_j| / r with r^2 = d^2 + eps2,
zero separations left out), so that a sink one ulp from a source, whose
field reaches 1e32 without softening, leaves every other sink checked at its
own scale.  The grouped mode is checked the same way against the oracle's
walk of the runs it forms."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = 24


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pepc_amd.solver import Solver
    return Solver


def _case(seed):
    rng = np.random.default_rng(3000 + seed)
    n = int(np.exp(rng.uniform(np.log(2.0), np.log(12000.0))))
    kind = int(rng.integers(0, 3))
    p = rng.random((n, 3))
    if kind == 1 and n >= 6:
        k = n // 3
        p[:k] = 0.4 + 1e-3 * rng.random((k, 3))
    if kind == 2 and n >= 20:
        m = n // 10
        idx = rng.permutation(n)
        p[idx[m:2 * m]] = p[idx[:m]]
    q = rng.choice(np.array([-1.0, 1.0]), n) if rng.random() < 0.5 else rng.normal(size=n)
    m = int(rng.integers(1, 4000))
    s = rng.uniform(-0.2, 1.2, (m, 3))                 # some outside the source box
    parts = rng.integers(0, 5, m)
    on = parts == 1                                    # exactly on a source
    s[on] = p[rng.integers(0, n, on.sum())]
    near = parts == 2                                  # a few ulps off a source
    s[near] = p[rng.integers(0, n, near.sum())]
    s[near] = np.nextafter(s[near], s[near] + rng.choice([-1.0, 1.0], (near.sum(), 3)))
    rep = np.flatnonzero(parts == 3)                   # repeats of other sinks
    if len(rep):
        s[rep] = s[rng.integers(0, m, len(rep))]
    clu = parts == 4                                   # a cluster of sinks
    s[clu] = 0.7 + 1e-4 * rng.random((clu.sum(), 3))
    theta2 = float(rng.choice([0.0, 0.09, 0.36, 0.81]))
    sign = float(rng.choice([1.0, -1.0]))
    eps2 = float(rng.choice([0.0, 1e-4]))
    return p, q, s, dict(theta2=theta2, sign=sign, eps2=eps2, kind=kind)


def _abs_sums(sx, sy, sz, x, y, z, q, eps2, chunk=256):
    m = len(sx)
    aq = np.abs(q)
    B, C = np.zeros(m), np.zeros(m)
    for a in range(0, m, chunk):
        b = min(m, a + chunk)
        d2 = (sx[a:b, None] - x) ** 2 + (sy[a:b, None] - y) ** 2 + (sz[a:b, None] - z) ** 2
        with np.errstate(divide="ignore"):
            ir = np.where(d2 == 0.0, 0.0, 1.0 / np.sqrt(d2 + eps2))
        C[a:b] = ir @ aq
        B[a:b] = (ir * ir) @ aq
    return B, C


@pytest.mark.parametrize("seed", range(CASES))
def test_random_sinks_parity(gpu, orc, seed):
    from oracle import periodic as P
    p, q, s, c = _case(seed)
    x, y, z = (p[:, k].copy() for k in range(3))
    sx, sy, sz = (s[:, k].copy() for k in range(3))
    print(f"[sinks seed {seed}] n={len(q)} m={len(sx)} {c}", flush=True)
    d = orc.domain_decompose(x, y, z, q)
    t = orc.tree_build(d.keys, d.x, d.y, d.z, d.q, d.box)
    w = P.walk_sinks(sx, sy, sz, t, len(q), c["theta2"], c["eps2"], sign=c["sign"])
    sol = gpu(theta2=c["theta2"], eps2=c["eps2"], precision="fp64", sign=c["sign"])
    sol.grow(x, y, z, q)
    res = sol.traverse_sinks(sx, sy, sz)
    p2p, m2p = sol.sink_counts(len(sx))
    assert np.array_equal(p2p.astype(np.int32), w.p2p)
    assert np.array_equal(m2p.astype(np.int32), w.m2p)
    assert np.array_equal(res.work.cpu().numpy(), w.work - 1.0)
    e = res.e.cpu().numpy().T
    pot = res.pot.cpu().numpy()
    B, C = _abs_sums(sx, sy, sz, d.x, d.y, d.z, d.q, c["eps2"])
    de = np.abs(e - w.e).max(axis=1)
    dp = np.abs(pot - w.pot)
    print(f"  max dE/B {np.max(de / np.maximum(B, 1e-300)):.1e} "
          f"max dphi/C {np.max(dp / np.maximum(C, 1e-300)):.1e}", flush=True)
    assert np.all(de <= 1e-12 * B) and np.all(dp <= 1e-12 * C)
    sol.close()


def _sink_groups(sx, sy, sz, box, G, orc):
    """The grouped mode's walk groups: sinks clamped into the source box,
    keyed on its Hilbert curve (20 levels, cell = size / 2^20), stably sorted,
    cut into runs of G.  Returns (order, tile_first, tile_count)."""
    bmin, bmax, bsize = box.as_arrays()
    ic = []
    for v, d in ((sx, 0), (sy, 1), (sz, 2)):
        w = np.minimum(np.maximum(v, bmin[d]), bmax[d])
        k = ((w - bmin[d]) / (bsize[d] / 1048576.0)).astype(np.int64)
        ic.append(np.clip(k, 0, (1 << 20) - 1))
    keys = np.array([orc.hilbert_key(a, b, c) for a, b, c in zip(*ic)], dtype=np.uint64)
    order = np.argsort(keys, kind="stable")
    m = len(sx)
    tf = np.arange(0, m, G, dtype=np.int64)
    tc = np.minimum(G, m - tf).astype(np.int64)
    return order, tf, tc


@pytest.mark.parametrize("seed", range(12))
def test_random_sinks_grouped_parity(gpu, orc, seed):
    """Grouped mode (no reference counterpart: sinks sorted along the source
    box's Hilbert curve, runs of group_size walking together) against the
    oracle's walk of the same runs: per-sink counts and work bit-exact, FP64
    fields and potentials to 1e-12 of each sink's own scale."""
    from oracle import periodic as P
    p, q, s, c = _case(seed)
    G = int(np.random.default_rng(7000 + seed).choice([2, 8, 16, 32, 64]))
    x, y, z = (p[:, k].copy() for k in range(3))
    sx, sy, sz = (s[:, k].copy() for k in range(3))
    print(f"[grouped sinks seed {seed}] n={len(q)} m={len(sx)} G={G} {c}", flush=True)
    d = orc.domain_decompose(x, y, z, q)
    t = orc.tree_build(d.keys, d.x, d.y, d.z, d.q, d.box)
    order, tf, tc = _sink_groups(sx, sy, sz, d.box, G, orc)
    w = P.walk_sinks(sx[order], sy[order], sz[order], t, len(q), c["theta2"], c["eps2"],
                     sign=c["sign"], tiles=(tf, tc))
    sol = gpu(theta2=c["theta2"], eps2=c["eps2"], precision="fp64", sign=c["sign"], group_size=G)
    sol.grow(x, y, z, q)
    res = sol.traverse_sinks(sx, sy, sz, grouped=True)
    p2p, m2p = sol.sink_counts(len(sx))
    assert np.array_equal(p2p[order].astype(np.int32), w.p2p)
    assert np.array_equal(m2p[order].astype(np.int32), w.m2p)
    assert np.array_equal(res.work.cpu().numpy()[order], w.work - 1.0)
    e = res.e.cpu().numpy().T[order]
    pot = res.pot.cpu().numpy()[order]
    B, C = _abs_sums(sx[order], sy[order], sz[order], d.x, d.y, d.z, d.q, c["eps2"])
    de = np.abs(e - w.e).max(axis=1)
    dp = np.abs(pot - w.pot)
    print(f"  max dE/B {np.max(de / np.maximum(B, 1e-300)):.1e} "
          f"max dphi/C {np.max(dp / np.maximum(C, 1e-300)):.1e}", flush=True)
    assert np.all(de <= 1e-12 * B) and np.all(dp <= 1e-12 * C)
    sol.close()
