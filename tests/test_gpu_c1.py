"""C1, the pepc-mini style case: 32,768 alternating +-1 charges in the unit
cube, eps = 0.01, theta = 0.6, quadrupole, FP64, one solve then 10 leapfrog
steps dt = 0.01 (unit masses, zero initial velocity), on the on-device
timestep.

At every step the GPU state (positions after that step's grow, in key
order) goes to the oracle's grow + group walk (module_tree_walk_simple rule,
TILE_SIZE = 16): the keys are already sorted (identity permutation on both
sides), interaction counts agree per target, fields agree to 1e-12, and the
step's energy record equals 1/2 sum q phi of the oracle to 1e-12.  Also:
pepc_gpu_restore_device returns traverse results to the input order.
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


def test_c1_ten_leapfrog_steps_per_step_parity(gpu, orc):
    s = workloads.mini()
    kw = dict(theta2=s.theta2, eps2=s.eps2, group_size=16, precision="fp64")
    stp = gpu(**kw)
    stp.state_set(s.x, s.y, s.z, s.v, s.q, s.m)
    ref = gpu(**kw)
    for step in range(11):
        st = stp.state_get()
        en = stp.energies()
        assert en.shape == (step + 1, 1, 2)
        d, t, w = orc.solve(st["x"], st["y"], st["z"], st["q"], s.theta2, s.eps2, tile_size=16)
        assert np.array_equal(d.perm, np.arange(s.n)), step
        ref.grow(st["x"], st["y"], st["z"], st["q"])
        res = ref.traverse()
        keys, perm, *_ = ref.sorted_host()
        assert np.array_equal(keys, d.keys) and np.array_equal(perm, np.arange(s.n))
        wk = ref.read_walk()
        assert np.array_equal(wk["p2p"].astype(np.int32), w.p2p)
        assert np.array_equal(wk["m2p"].astype(np.int32), w.m2p)
        e = res.e.cpu().numpy().T
        scale = np.abs(w.e).max()
        assert np.abs(e - w.e).max() <= 1e-12 * scale, step
        epot = 0.5 * np.sum(st["q"] * w.pot)
        assert abs(en[step, 0, 0] - epot) <= 1e-12 * abs(epot), step
        ekin = 0.5 * np.sum(st["m"] * np.sum(st["v"] ** 2, 1))
        assert abs(en[step, 0, 1] - ekin) <= 1e-12 * max(abs(ekin), 1e-300) + 1e-300, step
        print(f"[C1 step {step}] epot {en[step, 0, 0]:.12e} ekin {en[step, 0, 1]:.6e}", flush=True)
        if step < 10:
            stp.run(1, s.dt)
    stp.close()
    ref.close()


def test_restore_device(gpu):
    import torch
    s = workloads.mini(n=5000, seed=9)
    sol = gpu(theta2=s.theta2, eps2=s.eps2, precision="fp64")
    sol.grow(s.x, s.y, s.z, s.q)
    res = sol.traverse()
    keys, perm, xs, ys, zs, qs = sol.sorted_host()
    ex, pot, xr = sol.restore(res.e[0], res.pot, torch.from_numpy(xs))
    e_in = np.empty(s.n)
    e_in[perm] = res.e[0].cpu().numpy()
    assert np.array_equal(ex.cpu().numpy(), e_in)
    assert np.array_equal(xr.cpu().numpy(), s.x)     # no duplicate keys: x back exactly
    p_in = np.empty(s.n)
    p_in[perm] = res.pot.cpu().numpy()
    assert np.array_equal(pot.cpu().numpy(), p_in)
    sol.close()
