"""GPU adjoint (pepc_gpu_adjoint_device) against the oracle's adjoint, which
tests/test_oracle_adjoint.py pins by frozen-list finite differences and by the
closed-form pair-sum gradient.

* same tree, same tiles: gradients equal the oracle's to 1e-10 relative
  (different summation order and FP64 atomics only);
* theta = 0: equals the closed-form gradient of the pair sum;
* torch.autograd: coulomb_field's backward equals the adjoint, and
  gradcheck passes at theta = 0 (no acceptance decisions to freeze);
* batched systems: equals per-system adjoints.
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


def _charges(n, seed):
    r = np.random.default_rng(seed)
    p = r.random((n, 3))
    q = r.standard_normal(n)
    return r, p, q


@pytest.mark.parametrize("G,pack,theta2,sign", [(16, 0, 0.36, 1.0), (64, 64, 0.36, 1.0),
                                                (32, 0, 0.16, -1.0)])
def test_adjoint_matches_oracle(gpu, orc, G, pack, theta2, sign):
    import torch
    n, eps2 = 4096, 1e-4
    r, p, q = _charges(n, 61)
    sol = gpu(theta2=theta2, eps2=eps2, group_size=G, group_pack=pack, precision="fp64",
              sign=sign)
    sol.grow(p[:, 0], p[:, 1], p[:, 2], q)
    keys, perm, xs, ys, zs, qs = sol.sorted_host()
    ge = r.standard_normal((n, 3))
    gp = r.standard_normal(n)
    got = sol.adjoint(torch.from_numpy(ge.T.copy()), torch.from_numpy(gp))
    got = [g.cpu().numpy() for g in got]
    d = orc.domain_decompose(p[:, 0], p[:, 1], p[:, 2], q)
    assert np.array_equal(d.keys, keys)
    t = orc.tree_build(d.keys, d.x, d.y, d.z, d.q, d.box)
    tiles = orc.walk_tiles(d.keys, G, pack)
    lists = orc.walk_lists(d.x, d.y, d.z, t, theta2, tiles)
    want = orc.adjoint_lists(d.x, d.y, d.z, d.q, t, lists, eps2, ge, gp, sign=sign)
    print(f"[adjoint G={G} pack={pack}] entries={len(lists.node)}", flush=True)
    for name, a, b in zip("xyzq", got, want):
        scale = np.abs(b).max()
        err = np.abs(a - b).max() / scale
        print(f"  d/d{name}: max rel {err:.2e}", flush=True)
        assert err < 1e-10, name
    sol.close()


def test_adjoint_theta0_pair_sum(gpu):
    import torch
    n, eps2 = 500, 1e-3
    r, p, q = _charges(n, 62)
    sol = gpu(theta2=0.0, eps2=eps2, group_size=16, precision="fp64")
    sol.grow(p[:, 0], p[:, 1], p[:, 2], q)
    keys, perm, xs, ys, zs, qs = sol.sorted_host()
    ge = r.standard_normal((n, 3))
    gp = r.standard_normal(n)
    gx, gy, gz, gq = (g.cpu().numpy() for g in
                      sol.adjoint(torch.from_numpy(ge.T.copy()), torch.from_numpy(gp)))
    X = np.stack([xs, ys, zs], 1)
    D = X[:, None, :] - X[None, :, :]
    r2 = np.sum(D * D, 2) + eps2
    np.fill_diagonal(r2, 1.0)
    ri = 1.0 / np.sqrt(r2)
    np.fill_diagonal(ri, 0.0)
    r3, r5 = ri ** 3, ri ** 5
    ud = np.einsum("ik,ijk->ij", ge, D)
    T = qs[None, :, None] * (-gp[:, None, None] * D * r3[:, :, None] + ge[:, None, :] * r3[:, :, None]
                             - 3.0 * ud[:, :, None] * D * r5[:, :, None])
    want_x = T.sum(1) - T.sum(0)
    want_q = (gp[:, None] * ri + ud * r3).sum(0)
    got_x = np.stack([gx, gy, gz], 1)
    assert np.abs(got_x - want_x).max() <= 1e-11 * np.abs(want_x).max()
    assert np.abs(gq - want_q).max() <= 1e-11 * np.abs(want_q).max()
    sol.close()


def test_autograd_backward_and_gradcheck(gpu):
    import torch
    from pepc_amd.autograd import coulomb_field, energy
    dev = torch.device("cuda", 0)
    n = 3000
    r, p, q = _charges(n, 63)
    sol = gpu(theta2=0.36, eps2=1e-4, group_size=32, precision="fp64")
    X = [torch.tensor(p[:, k], device=dev, requires_grad=True) for k in range(3)]
    Q = torch.tensor(q, device=dev, requires_grad=True)
    e, pot = coulomb_field(*X, Q, sol)
    w = torch.randn(n, 3, dtype=torch.float64, device=dev)
    loss = (e * w).sum() + energy(Q, pot)
    loss.backward()
    # the same gradient through the C-ABI in key order
    keys, perm, xs, ys, zs, qs = sol.sorted_host()
    pt = torch.from_numpy(perm).to(dev)
    ge = w[pt].T.contiguous()
    gpot = 0.5 * Q.detach()[pt]
    gx, gy, gz, gq = sol.adjoint(ge, gpot)
    direct_q = torch.empty_like(gq)
    direct_q[pt] = gq
    direct_q += 0.5 * pot.detach()            # d(1/2 q.pot)/dq through q itself
    for g, a in zip((X[0].grad, X[1].grad, X[2].grad), (gx, gy, gz)):
        b = torch.empty_like(a)
        b[pt] = a
        assert torch.allclose(g, b, rtol=1e-12, atol=1e-12 * b.abs().max().item())
    assert torch.allclose(Q.grad, direct_q, rtol=1e-12, atol=1e-12 * direct_q.abs().max().item())
    sol.close()

    # gradcheck at theta = 0: the solve is the pair sum, smooth in every input
    m = 48
    r2, p2, q2 = _charges(m, 64)
    sol0 = gpu(theta2=0.0, eps2=1e-2, group_size=16, precision="fp64")
    inputs = tuple(torch.tensor(a, device=dev, requires_grad=True)
                   for a in (p2[:, 0], p2[:, 1], p2[:, 2], q2))

    def f(x, y, z, qq):
        e, pot = coulomb_field(x, y, z, qq, sol0)
        return e, pot

    # the adjoint sums its contributions with FP64 atomics from concurrent
    # waves, so two backward passes agree to rounding, not bit for bit
    assert torch.autograd.gradcheck(f, inputs, eps=1e-6, atol=1e-5, rtol=1e-4,
                                    nondet_tol=1e-10)
    sol0.close()


def test_batched_adjoint_equals_per_system(gpu):
    import torch
    systems = workloads.ensemble(nsys=5, seed=65, nmin=256, nmax=2048)
    r = np.random.default_rng(66)
    for s in systems:
        s.q[:] = r.standard_normal(s.n)
    x, y, z, q, m, v, off = workloads.concat(systems)
    n = int(off[-1])
    ge = r.standard_normal((3, n))
    gp = r.standard_normal(n)
    kw = dict(theta2=0.36, eps2=1e-4, group_size=32, precision="fp64")
    sol = gpu(**kw)
    sol.grow_batch(x, y, z, q, off)
    bg = [g.cpu().numpy() for g in sol.adjoint(torch.from_numpy(ge), torch.from_numpy(gp))]
    one = gpu(**kw)
    for k, s in enumerate(systems):
        a, b = off[k], off[k + 1]
        one.grow(s.x, s.y, s.z, s.q)
        g1 = [g.cpu().numpy() for g in one.adjoint(torch.from_numpy(ge[:, a:b].copy()),
                                                   torch.from_numpy(gp[a:b].copy()))]
        for u, w in zip(bg, g1):
            assert np.allclose(u[a:b], w, rtol=1e-12, atol=1e-12 * np.abs(w).max())
    sol.close()
    one.close()
