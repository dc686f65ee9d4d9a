"""Randomized timestep sweep: seeded systems (uniform, a dense cluster,
coincident pairs, whose duplicate-key fix-up rewrites positions at every
grow) with random velocities, masses, time step, precision, walk group rule
and interaction sign.  Three on-device kick-drift-kick steps equal, bit for
bit and step by step, their restatement on the host: the push arithmetic in
numpy (same operations, no contraction) around separate grow/traverse calls
that start from the previous step's order, as the device step does."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = 16
STEPS = 3


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pepc_amd.solver import Solver
    return Solver


def _case(seed):
    rng = np.random.default_rng(5000 + seed)
    n = int(np.exp(rng.uniform(np.log(2.0), np.log(6000.0))))
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
    v = rng.uniform(0.0, 0.5) * rng.standard_normal((n, 3))
    m = 0.5 + rng.random(n)
    G = int(rng.choice([1, 8, 16, 32, 64]))
    c = dict(theta2=float(rng.choice([0.0, 0.16, 0.36, 0.64])), eps2=float(rng.choice([1e-4, 1e-2])),
             group_size=G, group_pack=int(rng.choice([0, min(64, 2 * G)])) if 1 < G < 64 else 0,
             precision=str(rng.choice(["fp32", "fp64"])), sign=float(rng.choice([1.0, -1.0])))
    dt = float(rng.uniform(1e-4, 5e-3))
    return p, q, v, m, dt, c


def _solve(sol, x, y, z, q):
    sol.grow(x, y, z, q)
    res = sol.traverse()
    keys, perm, xs, ys, zs, qs = sol.sorted_host()
    return perm, xs, ys, zs, qs, res.e.cpu().numpy()


@pytest.mark.parametrize("seed", range(CASES))
def test_random_steps_match_host(gpu, seed):
    p, q, v, m, dt, c = _case(seed)
    n = len(q)
    print(f"[step seed {seed}] n={n} dt={dt:.2e} {c}", flush=True)
    # coincident pairs build deep twig chains; the device step does not regrow
    # an overflowing node pool, so both solvers get room for them
    stp = gpu(node_factor=8.0, **c)
    stp.state_set(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), v, q, m)
    ref = gpu(node_factor=8.0, **c)
    s = stp.state_get()
    perm, xs, ys, zs, qs, e = _solve(ref, p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), q)
    assert np.array_equal(s["label"], perm)
    assert np.array_equal(s["x"], xs) and np.array_equal(s["y"], ys) and np.array_equal(s["z"], zs)
    h = 0.5 * dt
    a = (s["q"] * e) / s["m"]              # (3, n), in the state's order
    for step in range(STEPS):
        vh = s["v"] + h * a.T
        x1 = [s[k] + dt * vh[:, j] for j, k in enumerate("xyz")]
        stp.run(1, dt)
        s1 = stp.state_get()
        perm1, xs1, ys1, zs1, qs1, e1 = _solve(ref, *x1, s["q"])
        m1 = s["m"][perm1]
        a = (qs1 * e1) / m1
        v1 = vh[perm1] + h * a.T
        tag = f"step {step + 1}"
        assert np.array_equal(s1["label"], s["label"][perm1]), tag
        assert np.array_equal(s1["x"], xs1) and np.array_equal(s1["y"], ys1), tag
        assert np.array_equal(s1["z"], zs1) and np.array_equal(s1["q"], qs1), tag
        assert np.array_equal(s1["m"], m1), tag
        assert np.array_equal(s1["v"], v1), tag
        s = s1
    stp.close()
    ref.close()
