"""Full-size workloads through size-independent properties (every-target
oracle parity at the benchmark sizes and rules is in
test_gpu_fullsize_oracle.py; these checks cover what the oracle does not
compute: the FP32 path against FP64, the direct-sum error, timesteps):

* C2 Coulomb explosion, N = 4,194,304, theta in {0.3, 0.6, 0.9}: keys sorted
  and in range, a permutation, tree and group invariants, work = 1 + counts,
  FP32 + Kahan within rounding of the FP64 tree, error against the FP64
  direct sum growing with theta, energy of 20 leapfrog steps;
* C3 clustered gravity, N = 8,388,608 (attractive sign): the same invariants,
  the sign identity (sign -1 gives exactly the negated sign +1 result), the
  direct-sum error;
* C4 256 batched systems: invariants per system, batch equals separate
  solves on sampled systems, autotuned graph timestep of 50 steps with
  per-system energy records;
* C5 adjoint, N = 262,144: the autograd gradient of
  L = sum |F - F_t|^2 + sum (phi - phi_t)^2 against FP64 central differences
  of the frozen-list replay on 4096 coordinates, and against the gradient of
  the same loss with the FP64 direct sum (theta = 0).
"""
import time

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


def _invariants(sol, st, res, x, y, z):
    keys, perm, xs, ys, zs, qs = sol.sorted_host()
    n = len(keys)
    assert np.all(keys[1:] > keys[:-1])
    assert keys[0] >= (1 << 60) and keys[-1] < (1 << 61)
    assert np.array_equal(np.sort(perm), np.arange(n))
    if st["dup_fixed"] == 0:
        assert np.array_equal(xs, x[perm]) and np.array_equal(zs, z[perm])
    assert st["nleaf"] == n and st["nnodes"] == st["nleaf"] + st["ntwig"]
    lo = np.array(st["level_offset"], np.int64)
    assert np.all(np.diff(lo[:st["maxlevel"] + 2]) >= 0) and lo[st["maxlevel"] + 1] == st["nnodes"]
    w = sol.read_walk()
    gf, gc = w["gfirst"].astype(np.int64), w["gcount"].astype(np.int64)
    assert gf[0] == 0 and np.array_equal(gf[1:], (gf + gc)[:-1]) and gc.sum() == n
    assert gc.max() <= 64
    p2p, m2p = w["p2p"].astype(np.int64), w["m2p"].astype(np.int64)
    assert p2p.min() >= 1                    # every target meets its own leaf
    assert w["stats"]["interactions"] == int(p2p.sum() + m2p.sum())
    assert np.array_equal(res.work.cpu().numpy(), 1.0 + p2p + m2p)
    return keys, perm, xs, ys, zs, qs


def _direct_err(sol, res, xs, ys, zs, qs, m=4096, seed=7):
    idx = workloads.sample_targets(len(xs), m, seed)
    ed, pd = sol.direct(xs, ys, zs, qs, idx)
    e = res.e[:, idx].T.cpu().numpy()
    ed = ed.cpu().numpy()
    rel = np.linalg.norm(e - ed, axis=1) / np.linalg.norm(ed, axis=1)
    return float(np.median(rel)), idx


def test_c2_explosion_full(gpu):
    import torch
    s = workloads.explosion()
    dev = torch.device("cuda", 0)
    X = [torch.from_numpy(a).to(dev) for a in (s.x, s.y, s.z, s.q)]
    meds = []
    for theta in (0.3, 0.6, 0.9):
        t0 = time.time()
        sol = gpu(theta2=theta * theta, eps2=s.eps2, group_size=64, group_pack=64)
        st = sol.grow(*X)
        res = sol.traverse()
        keys, perm, xs, ys, zs, qs = _invariants(sol, sol.sync(), res, s.x, s.y, s.z)
        med, idx = _direct_err(sol, res, xs, ys, zs, qs)
        meds.append(med)
        s64 = gpu(theta2=theta * theta, eps2=s.eps2, group_size=64, group_pack=64, precision="fp64")
        s64.grow(*X)
        r64 = s64.traverse()
        e32 = res.e[:, idx].T.cpu().numpy()
        e64 = r64.e[:, idx].T.cpu().numpy()
        rel = np.linalg.norm(e32 - e64, axis=1) / np.linalg.norm(e64, axis=1)
        print(f"[C2 theta={theta}] nodes={st['nnodes']} median err vs direct {med:.2e}, "
              f"fp32 vs fp64 tree median {np.median(rel):.2e} ({time.time() - t0:.1f} s)", flush=True)
        assert np.median(rel) < 2e-6
        sol.close()
        s64.close()
    assert meds[0] < meds[1] < meds[2]
    assert meds[1] < 1e-3
    # 20 leapfrog steps dt = 1e-3 (q = 1/N, m = 1), graph replay, energy records
    sol = gpu(theta2=0.36, eps2=s.eps2, group_size=64, group_pack=64)
    sol.state_set(s.x, s.y, s.z, s.v, s.q, s.m)
    sol.run(20, 1e-3, graph=True)
    en = sol.energies()
    tot = en[:, 0, 0] + en[:, 0, 1]
    drift = abs(tot[-1] - tot[0]) / abs(tot[0])
    print(f"[C2 20 steps] E0={tot[0]:.10e} drift {drift:.2e}", flush=True)
    assert en.shape == (21, 1, 2) and drift < 1e-6
    sol.close()


def test_c3_plummer_full(gpu):
    import torch
    s = workloads.plummer()
    dev = torch.device("cuda", 0)
    X = [torch.from_numpy(a).to(dev) for a in (s.x, s.y, s.z, s.q)]
    sol = gpu(theta2=s.theta2, eps2=s.eps2, sign=-1.0, group_size=64, group_pack=64)
    st = sol.grow(*X)
    res = sol.traverse()
    keys, perm, xs, ys, zs, qs = _invariants(sol, sol.sync(), res, s.x, s.y, s.z)
    med, idx = _direct_err(sol, res, xs, ys, zs, qs, m=2048)
    print(f"[C3] nodes={st['nnodes']} maxlevel={st['maxlevel']} dup_fixed={st['dup_fixed']} "
          f"median err vs direct {med:.2e}", flush=True)
    assert med < 1e-3
    pos = gpu(theta2=s.theta2, eps2=s.eps2, sign=1.0, group_size=64, group_pack=64)
    pos.grow(*X)
    rp = pos.traverse()
    assert torch.equal(res.e, -rp.e) and torch.equal(res.pot, -rp.pot)
    sol.close()
    pos.close()


def test_c4_ensemble_full(gpu):
    systems = workloads.ensemble()
    x, y, z, q, m, v, off = workloads.concat(systems)
    n = int(off[-1])
    sol = gpu(theta2=0.36, eps2=1e-4, precision="fp32")
    best, cands = sol.autotune(x, y, z, q, offsets=off, reps=1)
    print(f"[C4] n={n} nsys={len(systems)} autotuned {best}", flush=True)
    st = sol.grow_batch(x, y, z, q, off)
    res = sol.traverse()
    w = sol.read_walk()
    assert st["nleaf"] == n
    gf = w["gfirst"].astype(np.int64)
    gl = gf + w["gcount"].astype(np.int64) - 1
    assert np.array_equal(np.searchsorted(off, gf, "right"), np.searchsorted(off, gl, "right"))
    cfg = sol.config()
    one = gpu(theta2=0.36, eps2=1e-4, precision="fp32", group_size=cfg["group_size"],
              group_pack=cfg["group_pack"], list_tile=cfg["list_tile"])
    be, bp = res.e.cpu().numpy(), res.pot.cpu().numpy()
    for k in (0, 97, 255):
        a, b = off[k], off[k + 1]
        s = systems[k]
        one.grow(s.x, s.y, s.z, s.q)
        r1 = one.traverse()
        assert np.array_equal(be[:, a:b], r1.e.cpu().numpy())
        assert np.array_equal(bp[a:b], r1.pot.cpu().numpy())
    one.close()
    # 50 graph-replayed leapfrog steps, dt = 1e-3, energies per system
    sol.state_set(x, y, z, v, q, m, offsets=off)
    t0 = time.time()
    sol.run(50, 1e-3, graph=True)
    dt_run = time.time() - t0
    en = sol.energies()
    assert en.shape == (51, len(systems), 2) and np.all(np.isfinite(en))
    tot = en[:, :, 0] + en[:, :, 1]
    drift = np.abs(tot[-1] - tot[0]) / np.abs(en[0, :, 1])
    print(f"[C4] 50 steps in {dt_run:.2f} s; energy drift / E_kin: median {np.median(drift):.2e} "
          f"max {drift.max():.2e}", flush=True)
    lab = sol.state_get()["label"]
    for k in range(len(systems)):
        assert np.array_equal(np.sort(lab[off[k]:off[k + 1]]), np.arange(off[k], off[k + 1]))
    sol.close()

    # the batched timestep equals each system stepped alone, bit for bit
    pick = [0, 41, 255]
    sub = [systems[k] for k in pick]
    xs_, ys_, zs_, qs_, ms_, vs_, offs = workloads.concat(sub)
    kw = dict(theta2=0.36, eps2=1e-4, precision="fp32", group_size=cfg["group_size"],
              group_pack=cfg["group_pack"], list_tile=cfg["list_tile"])
    bat = gpu(**kw)
    bat.state_set(xs_, ys_, zs_, vs_, qs_, ms_, offsets=offs)
    bat.run(3, 1e-3, graph=True)
    bs = bat.state_get()
    for i, s_ in enumerate(sub):
        one = gpu(**kw)
        one.state_set(s_.x, s_.y, s_.z, s_.v, s_.q, s_.m)
        one.run(3, 1e-3)
        os_ = one.state_get()
        a, b = offs[i], offs[i + 1]
        for key in ("x", "y", "z", "v", "q", "m"):
            assert np.array_equal(bs[key][a:b], os_[key]), (i, key)
        assert np.array_equal(bs["label"][a:b] - a, os_["label"])
        one.close()
    bat.close()

    # the energy error is the leapfrog's (close +-1 pairs at eps = 0.01 are
    # barely resolved at dt = 1e-3): ten times smaller steps over the same
    # time shrink it
    sub = systems[:16]
    xs_, ys_, zs_, qs_, ms_, vs_, offs = workloads.concat(sub)
    drifts = []
    for dt, nst in ((1e-3, 50), (1e-4, 500)):
        b_ = gpu(**kw)
        b_.state_set(xs_, ys_, zs_, vs_, qs_, ms_, offsets=offs, max_records=nst + 1)
        b_.run(nst, dt, graph=True)
        e_ = b_.energies()
        t_ = e_[:, :, 0] + e_[:, :, 1]
        drifts.append(np.abs(t_[-1] - t_[0]) / np.abs(e_[0, :, 1]))
        b_.close()
    ratio = np.median(drifts[1]) / np.median(drifts[0])
    print(f"[C4] median drift dt=1e-3 {np.median(drifts[0]):.2e}, dt=1e-4 {np.median(drifts[1]):.2e}",
          flush=True)
    assert ratio < 0.2


def test_c5_adjoint_full(gpu):
    import torch
    from pepc_amd.autograd import coulomb_field
    s = workloads.adjoint()
    dev = torch.device("cuda", 0)
    n = s.n
    kw = dict(theta2=s.theta2, eps2=s.eps2, group_size=64, group_pack=64, precision="fp64")
    tgt = gpu(**kw)
    Xt = [torch.from_numpy(a).to(dev) for a in (s.x, s.y, s.z)]
    qt = torch.from_numpy(s.extra["q_target"]).to(dev)
    et, pt = coulomb_field(*Xt, qt, tgt)
    Ft = (qt[:, None] * et).detach()
    Pt = pt.detach()
    tgt.close()

    sol = gpu(**kw)
    X = [torch.from_numpy(a).to(dev).requires_grad_(True) for a in (s.x, s.y, s.z, s.q)]
    t0 = time.time()
    e, pot = coulomb_field(*X, sol)
    loss = ((X[3][:, None] * e - Ft) ** 2).sum() + ((pot - Pt) ** 2).sum()
    loss.backward()
    torch.cuda.synchronize()
    t_fb = time.time() - t0
    G = [t.grad.cpu().numpy() for t in X]

    # FD of the frozen-list replay (sorted order inside the solver)
    keys, perm, *_ = sol.sorted_host()
    pt_ = torch.from_numpy(perm).to(dev)
    Ft_s, Pt_s = Ft[pt_], Pt[pt_]
    inv = np.empty(n, np.int64)
    inv[perm] = np.arange(n)
    ne = sol.freeze()
    base = [t.detach().clone() for t in X]
    out = [None, None]

    def terms(Xs):
        o = sol.replay(*Xs, sync=False)
        A = Xs[3][pt_][None, :] * o.e            # (3, n) forces in sorted order
        return A.clone(), o.pot.clone()

    h = 1e-5
    rng = np.random.default_rng(91)
    coords = [(int(c), int(j)) for c, j in zip(rng.integers(0, 4, 4096), rng.integers(0, n, 4096))]
    fd = np.empty(len(coords))

    def diff(c, j, step):
        # L(+step) - L(-step), summed term by term (no cancellation of L itself)
        vals = []
        for sgn in (1.0, -1.0):
            Xs = list(base)
            Xs[c] = base[c].clone()
            Xs[c][j] += sgn * step
            vals.append(terms(Xs))
        (Ap, pp), (Am, pm) = vals
        return (((Ap - Am) * (Ap + Am - 2.0 * Ft_s.T)).sum()
                + ((pp - pm) * (pp + pm - 2.0 * Pt_s)).sum())

    t0 = time.time()
    for k, (c, j) in enumerate(coords):      # fourth-order central differences
        d = (8.0 * diff(c, j, h) - diff(c, j, 2.0 * h)) / (12.0 * h)
        fd[k] = d.item()
        if k % 512 == 0:
            print(f"  FD {k}/{len(coords)} ({time.time() - t0:.1f} s)", flush=True)
    t_fd = time.time() - t0
    ad = np.array([G[c][j] for c, j in coords])
    rms = np.array([np.sqrt(np.mean(G[c] ** 2)) for c, _ in coords])
    err = np.abs(fd - ad) / np.maximum(np.abs(ad), 1e-3 * rms)
    print(f"[C5] n={n} lists={ne} forward+backward {t_fb:.2f} s; FD over {len(coords)} coords "
          f"in {t_fd:.1f} s: median {np.median(err):.2e} max {err.max():.2e}", flush=True)
    assert np.median(err) < 1e-7 and err.max() < 1e-6

    # the same loss with the FP64 direct sum (theta = 0) for both the fields
    # and the targets: gradients agree to the tree's approximation error
    ds = gpu(theta2=0.0, eps2=s.eps2, group_size=64, group_pack=64, precision="fp64")
    t0 = time.time()
    etd, ptd = coulomb_field(*Xt, qt, ds)
    Ftd = (qt[:, None] * etd).detach()
    Ptd = ptd.detach()
    Xd = [torch.from_numpy(a).to(dev).requires_grad_(True) for a in (s.x, s.y, s.z, s.q)]
    e0, p0 = coulomb_field(*Xd, ds)
    l0 = ((Xd[3][:, None] * e0 - Ftd) ** 2).sum() + ((p0 - Ptd) ** 2).sum()
    l0.backward()
    torch.cuda.synchronize()
    Gd = [t.grad.cpu().numpy() for t in Xd]
    rels = [np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(G, Gd)]
    print(f"[C5] direct-sum gradient ({time.time() - t0:.1f} s): relative L2 difference "
          f"x {rels[0]:.2e} y {rels[1]:.2e} z {rels[2]:.2e} q {rels[3]:.2e}", flush=True)
    assert max(rels) < 0.1
    ds.close()
    sol.close()


def test_periodic_plasma_full(gpu):
    """Periodic plasma, N = 4,194,304 in the unit cube, 27 mirror boxes +
    lattice far field: the tree and group invariants, the device tree check,
    a neutral set of fictitious charges, and FP32 + Kahan within rounding of
    the FP64 periodic tree."""
    s = workloads.periodic(n=4194304, seed=71)
    out = {}
    for prec in ("fp64", "fp32"):
        sol = gpu(theta2=s.theta2, eps2=s.eps2, precision=prec, group_size=64, group_pack=64)
        sol.set_periodic()
        t0 = time.perf_counter()
        st = sol.grow(s.x, s.y, s.z, s.q)
        res = sol.traverse()
        dt = time.perf_counter() - t0
        st = sol.sync()
        _invariants(sol, st, res, s.x, s.y, s.z)
        assert sol.check_tree() == 0
        out[prec] = res.e.cpu().numpy().T
        print(f"[periodic 4M {prec}] {dt:.2f} s, {st['interactions'] / s.n:.0f} interactions/target",
              flush=True)
        if prec == "fp64":
            ml, mu, fict = sol.lattice()
            assert np.isfinite(mu).all() and len(fict) == 4
            assert abs(fict[:, 3].sum()) <= 1e-12 * max(1.0, np.abs(fict[:, 3]).max())
        sol.close()
    rel = np.linalg.norm(out["fp32"] - out["fp64"], axis=1) / np.linalg.norm(out["fp64"], axis=1)
    print(f"[periodic 4M] fp32 vs fp64 tree: median {np.median(rel):.2e} max {rel.max():.2e}",
          flush=True)
    assert np.median(rel) < 2e-6 and np.quantile(rel, 0.999) < 1e-4
