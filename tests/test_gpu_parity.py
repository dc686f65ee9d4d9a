"""GPU parity: the HIP path (through the C-ABI) against the oracle on the
same seeded inputs.  Integer work (keys, permutation, tree topology, groups,
interaction and MAC counts) must be bit-exact; FP64 moments are bit-exact
(same operation order, no contraction); forces agree to the stated
tolerances."""
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


SYSTEMS = {
    "mini": lambda: workloads.mini(),                      # 32768, +-1 charges
    "explosion64k": lambda: workloads.explosion(n=65536, seed=21),
    "plummer64k": lambda: workloads.plummer(n=65536, seed=22),
}

_cache = {}


def _oracle(orc, name, G, pack=0):
    key = (name, G, pack)
    if key not in _cache:
        s = SYSTEMS[name]()
        d = orc.domain_decompose(s.x, s.y, s.z, s.q)
        t = orc.tree_build(d.keys, d.x, d.y, d.z, d.q, d.box)
        w = orc.walk_group(d.x, d.y, d.z, t, s.theta2, s.eps2, tile_size=G, sign=s.sign,
                           keys=d.keys, pack=pack)
        _cache[key] = (s, d, t, w)
    return _cache[key]


def _run(gpu, s, G, precision, pack=0):
    sol = gpu(theta2=s.theta2, eps2=s.eps2, group_size=G, precision=precision, sign=s.sign,
              group_pack=pack)
    st = sol.grow(s.x, s.y, s.z, s.q)
    res = sol.traverse()
    return sol, st, res


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_keys_permutation_tree_bitexact(gpu, orc, name):
    s, d, t, _ = _oracle(orc, name, 16)
    sol, st, _ = _run(gpu, s, 16, "fp64")
    keys, perm, x, y, z, q = sol.sorted_host()
    print(f"[{name}] n={s.n} nodes={st['nnodes']} twigs={st['ntwig']}", flush=True)
    assert np.array_equal(keys, d.keys)
    assert np.array_equal(perm, d.perm)
    assert np.array_equal(x, d.x) and np.array_equal(y, d.y) and np.array_equal(z, d.z)
    assert st["nnodes"] == t.nnodes and st["ntwig"] == t.ntwig and st["nleaf"] == s.n
    tr = sol.read_tree()
    link, first, mom = tr["link"], tr["first"], tr["mom"]
    # match nodes by (level, first particle): unique per node
    og = np.lexsort((t.first, t.level))
    gg = np.lexsort((first, link[:, 3]))
    assert np.array_equal(t.level[og], link[gg, 3].astype(np.int32))
    assert np.array_equal(t.first[og], first[gg].astype(np.int64))
    assert np.array_equal(t.leaves[og], link[gg, 2].astype(np.int64))
    assert np.array_equal((t.first_child[og] < 0), link[gg, 1] == 0)
    assert np.array_equal(t.nchild[og], link[gg, 1].astype(np.int32))
    # FP64 moments, bit for bit (15 used slots)
    assert np.array_equal(t.mom[og][:, :15], mom[gg][:, :15])
    sol.close()


@pytest.mark.parametrize("G,pack", [(1, 0), (16, 0), (32, 0), (64, 0), (16, 32), (16, 64),
                                    (32, 32), (32, 64), (64, 64)])
@pytest.mark.parametrize("name", ["mini", "plummer64k"])
def test_groups_and_counts_bitexact(gpu, orc, name, G, pack):
    s, d, t, w = _oracle(orc, name, G, pack)
    sol, st, res = _run(gpu, s, G, "fp64", pack)
    wk = sol.read_walk()
    print(f"[{name} G={G}] groups={st['ngroups']} interactions={wk['stats']['interactions']}",
          flush=True)
    assert np.array_equal(wk["gfirst"].astype(np.int64), w.tile_first)
    assert np.array_equal(wk["gcount"].astype(np.int64), w.tile_count)
    assert np.array_equal(wk["p2p"].astype(np.int32), w.p2p)
    assert np.array_equal(wk["m2p"].astype(np.int32), w.m2p)
    assert np.array_equal(wk["gmac"].astype(np.int64), w.tile_mac)
    assert wk["stats"]["interactions"] == w.num_interactions
    assert wk["stats"]["mac_evaluations"] == w.num_mac
    assert np.array_equal(res.work.cpu().numpy(), w.work)
    sol.close()


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_fp64_forces_match_oracle(gpu, orc, name):
    s, d, t, w = _oracle(orc, name, 16)
    sol, st, res = _run(gpu, s, 16, "fp64")
    e = res.e.cpu().numpy().T
    pot = res.pot.cpu().numpy()
    scale = np.abs(w.e).max()
    # same terms, different summation order and rsqrt rounding: 1e-12 relative
    assert np.max(np.abs(e - w.e)) <= 1e-12 * scale
    assert np.max(np.abs(pot - w.pot)) <= 1e-12 * np.abs(w.pot).max()
    sol.close()


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_fp32_kahan_forces_close_to_fp64_tree(gpu, orc, name):
    s, d, t, w = _oracle(orc, name, 16)
    sol, st, res = _run(gpu, s, 16, "fp32")
    e = res.e.cpu().numpy().T
    rel = np.linalg.norm(e - w.e, axis=1) / np.linalg.norm(w.e, axis=1)
    prel = np.abs(res.pot.cpu().numpy() - w.pot) / np.abs(w.pot).max()
    print(f"[{name}] fp32 vs fp64 tree: median {np.median(rel):.2e} max {rel.max():.2e}", flush=True)
    # FP32 arithmetic on identical interaction lists: rounding level only
    assert np.median(rel) < 2e-6
    assert np.quantile(rel, 0.999) < 1e-4
    assert prel.max() < 1e-5
    wk = sol.read_walk()
    assert np.array_equal(wk["p2p"].astype(np.int32), w.p2p)
    sol.close()


def test_direct_sum_matches_oracle(gpu, orc):
    s = workloads.mini(n=4096, seed=31)
    sol = gpu(theta2=0.36, eps2=s.eps2)
    idx = np.array([0, 5, 77, 4095])
    e, pot = sol.direct(s.x, s.y, s.z, s.q, idx)
    oe, op = orc.direct(s.x, s.y, s.z, s.q, idx, s.eps2)
    assert np.allclose(e.cpu().numpy(), oe, rtol=1e-12, atol=1e-12 * np.abs(oe).max())
    assert np.allclose(pot.cpu().numpy(), op, rtol=1e-12)
    sol.close()


def test_coincident_particles_dupfix(gpu, orc):
    # pairs of particles at identical positions exercise the duplicate-key
    # fix-up (module_domains.f90:248-307); opposite charges on the pairs, so
    # which particle of a tied pair moves shows in the charges and fields
    r = np.random.default_rng(5)
    p = r.random((3000, 3))
    p[1000:1010] = p[0:10]
    q = np.ones(3000)
    q[1000:1010] = -1.0
    d = orc.domain_decompose(p[:, 0], p[:, 1], p[:, 2], q)
    assert d.nfixed > 0
    sol = gpu(theta2=0.36, eps2=1e-4, precision="fp64")
    st = sol.grow(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), q)
    keys, perm, x, y, z, qs = sol.sorted_host()
    assert st["dup_fixed"] == 10
    assert np.array_equal(keys, d.keys)
    assert np.array_equal(perm, d.perm)
    assert np.array_equal(x, d.x) and np.array_equal(y, d.y) and np.array_equal(z, d.z)
    assert np.array_equal(qs, d.q)
    t = orc.tree_build(d.keys, d.x, d.y, d.z, d.q, d.box)
    assert st["nnodes"] == t.nnodes
    w = orc.walk_group(d.x, d.y, d.z, t, 0.36, 1e-4, tile_size=16, keys=d.keys)
    res = sol.traverse()
    assert np.max(np.abs(res.e.cpu().numpy().T - w.e)) <= 1e-12 * np.abs(w.e).max()
    sol.close()


def test_error_paths(gpu):
    from pepc_amd import _lib
    sol = gpu()
    x = np.array([0.0, np.nan, 1.0])
    with pytest.raises(_lib.PepcGpuError) as ei:
        sol.grow(x, x.copy(), x.copy(), np.ones(3))
    assert ei.value.code == _lib.EBOX
    sol.close()
    sol = gpu(node_factor=1.0)   # too small a node pool for a real tree
    s = workloads.mini(n=20000)
    sol.grow(s.x, s.y, s.z, s.q, sync=False)   # the device form reports it at the next sync
    with pytest.raises(_lib.PepcGpuError) as ei:
        sol.sync()
    assert ei.value.code == _lib.EOVERFLOW
    sol.close()
    # a synchronous grow enlarges the pool to the counted size and grows again
    sol = gpu(node_factor=1.0)
    st = sol.grow(s.x, s.y, s.z, s.q)
    assert st["nleaf"] == s.n and st["nnodes"] > 1.0 * (s.n + s.n // 8 + 64)
    sol.close()
    # a node pool beyond 31-bit node indices (the walk keeps the node kind in
    # the top bit of its stack entries) is refused before any allocation,
    # also between 2^31 and 2^32 nodes
    for nf in (1e6, 1.2e5):
        sol = gpu(node_factor=nf)
        with pytest.raises(_lib.PepcGpuError) as ei:
            sol.grow(s.x, s.y, s.z, s.q)
        assert ei.value.code == _lib.EINVAL and "31-bit" in str(ei.value)
        sol.close()


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_tree_check_sound(gpu, name):
    """pepc_gpu_check_tree (tree_check behind pepc_check_sanity) finds the
    grown trees sound, also for a batch with an empty system."""
    s = SYSTEMS[name]()
    sol = gpu(theta2=s.theta2, eps2=s.eps2, precision="fp32", sign=s.sign)
    sol.grow(s.x, s.y, s.z, s.q)
    assert sol.check_tree() == 0
    k = s.n // 3
    off = np.array([0, k, k, s.n], np.int64)
    sol.grow_batch(s.x, s.y, s.z, s.q, off)
    assert sol.check_tree() == 0
    sol.close()


def test_direct_host_form(gpu, orc):
    """pepc_gpu_direct (host arrays) equals the oracle's direct sum and
    rejects test indices outside [0, n) with EINVAL before any device work."""
    import ctypes as C
    from pepc_amd import _lib
    s = workloads.mini(n=2048, seed=5)
    sol = gpu(theta2=s.theta2, eps2=s.eps2)
    src = [np.ascontiguousarray(a, dtype=np.float64) for a in (s.x, s.y, s.z, s.q)]
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    for idx, want in ((np.array([0, 7, 2047, 1000], np.int64), _lib.OK),
                      (np.array([0, 2048], np.int64), _lib.EINVAL),
                      (np.array([-1], np.int64), _lib.EINVAL)):
        out = [np.zeros(len(idx)) for _ in range(4)]
        rc = sol.lib.pepc_gpu_direct(sol.h, s.n, *[vp(a) for a in src], len(idx), vp(idx),
                                     s.eps2, 1.0, *[vp(a) for a in out])
        assert rc == want
        if rc == _lib.OK:
            e_ref, p_ref = orc.direct(s.x, s.y, s.z, s.q, idx, s.eps2)
            e = np.stack(out[:3], 1)
            assert np.max(np.abs(e - e_ref)) <= 1e-12 * np.abs(e_ref).max()
            assert np.max(np.abs(out[3] - p_ref)) <= 1e-12 * np.abs(p_ref).max()
    sol.close()


def test_host_grow_enlarges_the_node_pool(gpu, orc):
    """The host form pepc_gpu_grow (the Fortran shim's path) never fails for
    lack of nodes: with a pool of one node per particle, too small for any
    real tree, it grows the pool to the count the build found and returns
    the oracle's keys and permutation; pepc_gpu_reserve_nodes raises the pool
    ahead of time."""
    import ctypes as C
    from pepc_amd import _lib
    s = workloads.mini(n=20000, seed=9)
    d = orc.domain_decompose(s.x, s.y, s.z, s.q)
    t = orc.tree_build(d.keys, d.x, d.y, d.z, d.q, d.box)
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    src = [np.ascontiguousarray(a, dtype=np.float64) for a in (s.x, s.y, s.z, s.q)]
    for reserve in (0, 3 * s.n):
        sol = gpu(theta2=s.theta2, eps2=s.eps2, node_factor=1.0)
        if reserve:
            assert sol.lib.pepc_gpu_reserve_nodes(sol.h, reserve) == _lib.OK
        keys = np.zeros(s.n, np.uint64)
        perm = np.zeros(s.n, np.int64)
        xs, ys, zs = (np.zeros(s.n) for _ in range(3))
        rc = sol.lib.pepc_gpu_grow(sol.h, s.n, *[vp(a) for a in src], vp(keys), vp(perm),
                                   vp(xs), vp(ys), vp(zs))
        assert rc == _lib.OK
        st = sol.sync()
        assert st["nnodes"] == t.nnodes
        assert np.array_equal(keys, d.keys) and np.array_equal(perm, d.perm)
        assert np.array_equal(xs, d.x)
        assert sol.check_tree() == 0
        sol.close()
    sol = gpu()
    assert sol.lib.pepc_gpu_reserve_nodes(sol.h, -1) == _lib.EINVAL
    sol.close()
