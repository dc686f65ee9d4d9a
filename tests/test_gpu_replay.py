"""Frozen interaction lists on the GPU (pepc_gpu_freeze / pepc_gpu_replay_device).

* the recorded lists have one entry per (group, interaction node): their
  total times the group sizes equals the walk's interaction count;
* replaying at the grown positions reproduces the FP64 walk (summation order
  only);
* central differences of the replay equal the GPU adjoint (h = 1e-5,
  fourth order), which is the adjoint's definition: the derivative of the
  solve with lists held fixed.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pepc_amd.solver import Solver
    return Solver


def test_replay_reproduces_walk_and_adjoint_matches_differences(gpu):
    import torch
    n, eps2 = 3000, 1e-4
    r = np.random.default_rng(81)
    p = r.random((n, 3))
    q = r.standard_normal(n)
    sol = gpu(theta2=0.36, eps2=eps2, group_size=32, group_pack=64, precision="fp64")
    sol.grow(p[:, 0], p[:, 1], p[:, 2], q)
    res = sol.traverse()
    w = sol.read_walk()
    ne = sol.freeze()
    st = w["stats"]
    # each list entry is one interaction for every target of its group
    assert ne > 0 and st["interactions"] > ne
    X = [torch.tensor(a, device="cuda") for a in (p[:, 0], p[:, 1], p[:, 2], q)]
    rep = sol.replay(*X)
    e0, p0 = res.e.cpu().numpy(), res.pot.cpu().numpy()
    e1, p1 = rep.e.cpu().numpy(), rep.pot.cpu().numpy()
    assert np.abs(e1 - e0).max() <= 1e-12 * np.abs(e0).max()
    assert np.abs(p1 - p0).max() <= 1e-12 * np.abs(p0).max()

    keys, perm, *_ = sol.sorted_host()
    ge = r.standard_normal((3, n))
    gp = r.standard_normal(n)
    sol.replay(*X)        # positions back to the grown ones
    G = [g.cpu().numpy() for g in sol.adjoint(torch.from_numpy(ge), torch.from_numpy(gp))]
    inv = np.empty(n, np.int64)
    inv[perm] = np.arange(n)
    ge_t = torch.from_numpy(ge).cuda()
    gp_t = torch.from_numpy(gp).cuda()

    def loss(Xs):
        o = sol.replay(*Xs)
        return float((o.e * ge_t).sum() + (o.pot * gp_t).sum())

    h = 1e-5
    worst = 0.0
    for k in range(64):
        c, j = k % 4, int(r.integers(n))     # coordinate c of input particle j

        def f(s):
            Xs = [a.clone() for a in X]
            Xs[c][j] += s
            return loss(Xs)

        fd = (8 * (f(h) - f(-h)) - (f(2 * h) - f(-2 * h))) / (12 * h)
        a = G[c][inv[j]]
        scale = max(abs(a), 1e-3 * np.sqrt(np.mean(G[c] ** 2)))
        worst = max(worst, abs(fd - a) / scale)
    print(f"worst relative FD deviation {worst:.2e}", flush=True)
    assert worst <= 1e-6
    sol.close()
