"""Randomized adjoint sweep: seeded systems (uniform, a dense cluster,
coincident pairs) with random opening angle, walk group rule, sign,
softening and upstream gradients; the GPU adjoint (interaction lists frozen)
against the oracle's adjoint over the oracle's own lists.  Each particle's
gradient agrees to 1e-10 of its own sum of absolute pair terms, as a target
and as a source (r^2 = d^2 + eps2):
  d/dx_i: sum_j |q_j| (3 |ge_i| / r^3 + |gp_i| / r^2)
                + |q_i| (3 |ge_j| / r^3 + |gp_j| / r^2)
  d/dq_i: sum_j |ge_j| / r^2 + |gp_j| / r
so that a coincident pair without softening, whose terms reach 1e24, leaves
every other particle checked at its own scale."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = 16


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pepc_amd.solver import Solver
    return Solver


def _case(seed):
    rng = np.random.default_rng(4000 + seed)
    n = int(np.exp(rng.uniform(np.log(2.0), np.log(8000.0))))
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
    ge = rng.normal(size=(n, 3))
    gp = rng.normal(size=n)
    theta2 = float(rng.choice([0.0, 0.09, 0.36, 0.81]))
    G = int(rng.choice([1, 8, 16, 32, 64]))
    pack = int(rng.choice([0, min(64, 2 * G)])) if 1 < G < 64 else 0
    sign = float(rng.choice([1.0, -1.0]))
    eps2 = float(rng.choice([0.0, 1e-4]))
    return p, q, ge, gp, dict(theta2=theta2, G=G, pack=pack, sign=sign, eps2=eps2, kind=kind)


def _scales(x, y, z, q, ge, gp, eps2, chunk=256):
    n = len(x)
    aq, age, agp = np.abs(q), np.abs(ge).max(axis=1), np.abs(gp)
    sx, sq = np.zeros(n), np.zeros(n)
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        d2 = (x[a:b, None] - x) ** 2 + (y[a:b, None] - y) ** 2 + (z[a:b, None] - z) ** 2
        with np.errstate(divide="ignore"):
            ir = np.where(d2 == 0.0, 0.0, 1.0 / np.sqrt(d2 + eps2))
        ir2 = ir * ir
        ir3 = ir2 * ir
        sx[a:b] = (aq[a:b] * ((3.0 * ir3) @ age + ir2 @ agp)
                   + (ir3 @ aq) * 3.0 * age[a:b] + (ir2 @ aq) * agp[a:b])
        sq[a:b] = ir2 @ age + ir @ agp
    return sx, sq


@pytest.mark.parametrize("seed", range(CASES))
def test_random_adjoint_parity(gpu, orc, seed):
    import torch
    p, q, ge, gp, c = _case(seed)
    n = len(q)
    print(f"[adjoint seed {seed}] n={n} {c}", flush=True)
    sol = gpu(theta2=c["theta2"], eps2=c["eps2"], group_size=c["G"], group_pack=c["pack"],
              precision="fp64", sign=c["sign"])
    sol.grow(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), q)
    keys, perm, xs, ys, zs, qs = sol.sorted_host()
    got = sol.adjoint(torch.from_numpy(ge.T.copy()), torch.from_numpy(gp))
    got = [g.cpu().numpy() for g in got]
    d = orc.domain_decompose(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), q)
    assert np.array_equal(d.keys, keys) and np.array_equal(d.x, xs)
    t = orc.tree_build(d.keys, d.x, d.y, d.z, d.q, d.box)
    tiles = orc.walk_tiles(d.keys, c["G"], c["pack"])
    lists = orc.walk_lists(d.x, d.y, d.z, t, c["theta2"], tiles)
    want = orc.adjoint_lists(d.x, d.y, d.z, d.q, t, lists, c["eps2"], ge, gp, sign=c["sign"])
    sx, sq = _scales(d.x, d.y, d.z, d.q, ge, gp, c["eps2"])
    for name, a, b in zip("xyzq", got, want):
        s = sq if name == "q" else sx
        ratio = np.abs(a - b) / np.maximum(s, 1e-300)
        print(f"  d/d{name}: max |diff| / scale {ratio.max():.1e}", flush=True)
        assert np.all(np.abs(a - b) <= 1e-10 * s), name
    sol.close()
