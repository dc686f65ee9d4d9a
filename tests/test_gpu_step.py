"""On-device kick-drift-kick timestep and its graph replay.

* one step equals its restatement on the host: the push arithmetic in numpy
  (same operations, no contraction) around separate grow/traverse calls,
  bit for bit;
* graph replay equals eager steps bit for bit on every replay, with eager
  work and synchronisations between replays;
* at theta = 0 the tree solve is the direct sum, so a trajectory matches a
  numpy leapfrog driven by the oracle's FP64 direct sum (1e-9).
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


def _solve(sol, x, y, z, q):
    sol.grow(x, y, z, q)
    res = sol.traverse()
    keys, perm, xs, ys, zs, qs = sol.sorted_host()
    return perm, xs, ys, zs, qs, res.e.cpu().numpy(), res.pot.cpu().numpy()


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_step_matches_host_restatement(gpu, precision):
    s = workloads.mini(n=4096, seed=51)
    r = np.random.default_rng(52)
    v = 0.1 * r.standard_normal((s.n, 3))
    m = 1.0 + r.random(s.n)
    dt = 0.003
    kw = dict(theta2=s.theta2, eps2=s.eps2, precision=precision, group_size=32)
    stp = gpu(**kw)
    stp.state_set(s.x, s.y, s.z, v, s.q, m)
    s0 = stp.state_get()
    ref = gpu(**kw)
    # initial solve: state in key order, accelerations from the same tree
    perm, xs, ys, zs, qs, e0, pot0 = _solve(ref, s.x, s.y, s.z, s.q)
    assert np.array_equal(s0["label"], perm)
    assert np.array_equal(s0["x"], xs) and np.array_equal(s0["q"], qs)
    assert np.array_equal(s0["v"], v[perm]) and np.array_equal(s0["m"], m[perm])
    en = stp.energies()
    assert en.shape == (1, 1, 2)
    assert np.isclose(en[0, 0, 0], 0.5 * np.sum(qs * pot0), rtol=1e-12)
    assert np.isclose(en[0, 0, 1], 0.5 * np.sum(m * np.sum(v * v, 1)), rtol=1e-12)

    a0 = (s0["q"] * e0) / s0["m"]          # (3, n)
    h = 0.5 * dt
    vh = s0["v"] + h * a0.T
    x1 = [s0[c] + dt * vh[:, k] for k, c in enumerate("xyz")]
    stp.run(1, dt)
    s1 = stp.state_get()
    perm1, xs1, ys1, zs1, qs1, e1, pot1 = _solve(ref, *x1, s0["q"])
    m1 = s0["m"][perm1]
    a1 = (qs1 * e1) / m1
    v1 = vh[perm1] + h * a1.T
    assert np.array_equal(s1["x"], xs1) and np.array_equal(s1["y"], ys1)
    assert np.array_equal(s1["z"], zs1) and np.array_equal(s1["q"], qs1)
    assert np.array_equal(s1["m"], m1)
    assert np.array_equal(s1["label"], s0["label"][perm1])
    assert np.array_equal(s1["v"], v1)
    en = stp.energies()
    assert en.shape == (2, 1, 2)
    assert np.isclose(en[1, 0, 0], 0.5 * np.sum(qs1 * pot1), rtol=1e-12)
    assert np.isclose(en[1, 0, 1], 0.5 * np.sum(m1 * np.sum(v1 * v1, 1)), rtol=1e-12)
    stp.close()
    ref.close()


def test_graph_replay_matches_eager(gpu):
    import torch
    systems = workloads.ensemble(nsys=6, seed=53, nmin=256, nmax=4096)
    x, y, z, q, m, v, off = workloads.concat(systems)
    kw = dict(theta2=0.36, eps2=1e-4, precision="fp32", group_size=64, group_pack=64)
    dt = 1e-3
    eager = gpu(**kw)
    eager.state_set(x, y, z, v, q, m, offsets=off)
    graph = gpu(**kw)
    graph.state_set(x, y, z, v, q, m, offsets=off)
    junk = torch.randn(1 << 16, device="cuda")
    for step in range(4):
        eager.run(1, dt)
        graph.run(1, dt, graph=True)
        a, b = eager.state_get(), graph.state_get()
        print(f"[graph replay {step}] n={off[-1]}", flush=True)
        for key in a:
            assert np.array_equal(a[key], b[key]), (step, key)
        junk = (junk * 1.0001).sum() + junk     # eager work between replays
        torch.cuda.synchronize()
    # several replays in one call
    eager.run(3, dt)
    graph.run(3, dt, graph=True)
    a, b = eager.state_get(), graph.state_get()
    for key in a:
        assert np.array_equal(a[key], b[key]), key
    ea, eb = eager.energies(), graph.energies()
    assert ea.shape == (8, len(systems), 2)
    assert np.array_equal(ea, eb)
    # labels stay a permutation within each system's range
    for k in range(len(systems)):
        lab = a["label"][off[k]:off[k + 1]]
        assert np.array_equal(np.sort(lab), np.arange(off[k], off[k + 1]))
    eager.close()
    graph.close()


def test_direct_sum_trajectory(gpu, orc):
    r = np.random.default_rng(54)
    n = 400
    p = r.random((n, 3))
    q = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    v = 0.1 * r.standard_normal((n, 3))
    m = np.ones(n)
    eps2, dt, nsteps = 1e-2, 2e-3, 10
    sol = gpu(theta2=0.0, eps2=eps2, precision="fp64", group_size=16)
    sol.state_set(p[:, 0], p[:, 1], p[:, 2], v, q, m)
    sol.run(nsteps, dt, graph=True)
    got = sol.state_get()

    idx = np.arange(n)

    def accel(pos):
        e, pot = orc.direct(pos[:, 0].copy(), pos[:, 1].copy(), pos[:, 2].copy(), q, idx, eps2)
        return (q[:, None] * e) / m[:, None], pot

    pos, vel = p.copy(), v.copy()
    a, pot = accel(pos)
    e_hist = [0.5 * np.sum(q * pot)]
    for _ in range(nsteps):
        vel = vel + 0.5 * dt * a
        pos = pos + dt * vel
        a, pot = accel(pos)
        vel = vel + 0.5 * dt * a
        e_hist.append(0.5 * np.sum(q * pot))
    lab = got["label"]
    gpos = np.stack([got["x"], got["y"], got["z"]], 1)
    assert np.max(np.abs(gpos - pos[lab])) < 1e-9
    assert np.max(np.abs(got["v"] - vel[lab])) < 1e-8
    en = sol.energies()
    assert en.shape == (nsteps + 1, 1, 2)
    assert np.allclose(en[:, 0, 0], e_hist, rtol=1e-10)
    sol.close()
