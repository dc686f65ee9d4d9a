"""Pins the oracle's kernels (coulomb/module_coulomb_kernels.f90) and group
walk (module_tree_walk_simple.f90) by closed forms, theta=0 reduction to the
direct sum, an independent per-particle walk, and the FP64 direct sum."""
import numpy as np
import pytest

from pepc_amd import workloads


def test_p2p_closed_form(orc):
    d = np.array([0.3, -0.2, 0.7]); q = -1.7; eps2 = 1e-3
    e, phi = orc.p2p(q, d, d @ d + eps2)
    r = np.sqrt(d @ d + eps2)
    assert np.allclose(e, q * d / r ** 3, rtol=1e-15)
    assert np.isclose(phi, q / r, rtol=1e-15)


def _mom(ch, dip=(0, 0, 0), quad=(0, 0, 0), off=(0, 0, 0)):
    m = np.zeros(16)
    m[3] = ch; m[4] = abs(ch); m[5:8] = dip; m[8:11] = quad; m[11:14] = off
    return m


def test_m2p_monopole_equals_p2p(orc):
    d = np.array([1.1, 0.4, -0.9])
    e1, p1 = orc.m2p(_mom(2.5), d, d @ d + 0.01)
    e2, p2 = orc.p2p(2.5, d, d @ d + 0.01)
    assert np.allclose(e1, e2, rtol=1e-14) and np.isclose(p1, p2, rtol=1e-14)


def test_m2p_field_is_minus_gradient_of_potential(orc):
    r = np.random.default_rng(0)
    m = _mom(0.7, r.standard_normal(3), r.standard_normal(3), r.standard_normal(3))
    eps2 = 0.05
    d = np.array([0.8, -0.6, 1.3])
    e, _ = orc.m2p(m, d, d @ d + eps2)
    h = 1e-5
    g = np.zeros(3)
    for k in range(3):
        dp = d.copy(); dp[k] += h
        dm = d.copy(); dm[k] -= h
        g[k] = (orc.m2p(m, dp, dp @ dp + eps2)[1] - orc.m2p(m, dm, dm @ dm + eps2)[1]) / (2 * h)
    assert np.allclose(e, -g, rtol=1e-7, atol=1e-9)


def test_m2p_converges_to_cluster_direct_sum(orc):
    # expansion of a positive cluster about its centre of charge (dipole = 0):
    # with the quadrupole the relative field error falls like (size/dist)^3
    # (ratio ~8 per doubling); dropping the quadrupole leaves (size/dist)^2 (~4)
    r = np.random.default_rng(3)
    X = r.random((30, 3)) * 0.1
    q = r.random(30) + 0.2
    t = orc.tree_build(*_clusterkeys(orc, X, q))
    m = t.mom[0]
    errs, errs0 = [], []
    for dist in (1.0, 2.0, 4.0):
        xt = m[:3] + np.array([dist, 0.3 * dist, -0.2 * dist])
        dd = xt[None, :] - X
        r3 = ((dd ** 2).sum(1)) ** 1.5
        exact = (q[:, None] * dd / r3[:, None]).sum(0)
        e, _ = orc.m2p(m, xt - m[:3], (xt - m[:3]) @ (xt - m[:3]))
        errs.append(np.linalg.norm(e - exact) / np.linalg.norm(exact))
        m0 = m.copy(); m0[8:14] = 0
        e0, _ = orc.m2p(m0, xt - m[:3], (xt - m[:3]) @ (xt - m[:3]))
        errs0.append(np.linalg.norm(e0 - exact) / np.linalg.norm(exact))
    assert errs[0] < 1e-3
    assert errs[1] / errs[2] > 6.5 and errs[0] / errs[1] > 6.5
    assert errs0[1] / errs0[2] < 5 and all(a < b for a, b in zip(errs, errs0))


def _clusterkeys(orc, X, q):
    d = orc.domain_decompose(X[:, 0], X[:, 1], X[:, 2], q)
    return d.keys, d.x, d.y, d.z, d.q, d.box


def test_theta0_is_direct_sum(orc):
    s = workloads.mini(n=1500, seed=9)
    d, t, w = orc.solve(s.x, s.y, s.z, s.q, theta2=0.0, eps2=s.eps2)
    assert np.all(w.p2p == 1500) and np.all(w.m2p == 0)
    assert np.all(w.work == 1 + 1500)
    e, pot = orc.direct(d.x, d.y, d.z, d.q, np.arange(1500), s.eps2)
    assert np.allclose(w.e, e, rtol=1e-11, atol=1e-11 * np.abs(e).max())
    assert np.allclose(w.pot, pot, rtol=1e-11)


def test_direct_sum_textbook(orc):
    s = workloads.mini(n=400, seed=4)
    idx = np.array([0, 17, 399])
    e, pot = orc.direct(s.x, s.y, s.z, s.q, idx, s.eps2)
    P = np.stack([s.x, s.y, s.z], 1)
    for j, t in enumerate(idx):
        dd = P[t] - P
        r2 = (dd ** 2).sum(1) + s.eps2
        msk = np.arange(400) != t
        assert np.allclose(e[j], (s.q[msk, None] * dd[msk] / r2[msk, None] ** 1.5).sum(0), rtol=1e-12)
        assert np.isclose(pot[j], (s.q[msk] / np.sqrt(r2[msk])).sum(), rtol=1e-12)


def _py_walk_particle(t, d, theta2, eps2, orc, ip):
    """Independent per-particle walk (pthreads walk semantics,
    module_tree_walk_pthreads.f90:604-649): explicit stack, leaf -> P2P,
    twig -> MAC -> M2P or open."""
    x = np.array([d.x[ip], d.y[ip], d.z[ip]])
    e = np.zeros(3); pot = 0.0; np2 = 0; nm2 = 0
    stack = [0]
    while stack:
        nd = stack.pop()
        m = t.mom[nd]
        dd = x - m[:3]
        d2 = dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]
        if t.first_child[nd] < 0:
            np2 += 1
            if d2 > 0:
                ee, pp = orc.p2p(m[3], dd, d2 + eps2); e += ee; pot += pp
        elif theta2 * d2 > t.b2[t.level[nd]]:
            nm2 += 1
            ee, pp = orc.m2p(m, dd, d2 + eps2); e += ee; pot += pp
        else:
            c = t.first_child[nd]
            kids = []
            while c >= 0:
                kids.append(c); c = t.next_sibling[c]
            stack.extend(kids)
    return e, pot, np2, nm2


def test_g1_walk_equals_per_particle_walk(orc):
    s = workloads.mini(n=600, seed=6)
    d, t, w = orc.solve(s.x, s.y, s.z, s.q, theta2=s.theta2, eps2=s.eps2, tile_size=1)
    assert len(w.tile_first) == 600
    for ip in range(0, 600, 37):
        e, pot, np2, nm2 = _py_walk_particle(t, d, s.theta2, s.eps2, orc, ip)
        assert w.p2p[ip] == np2 and w.m2p[ip] == nm2
        assert np.allclose(w.e[ip], e, rtol=1e-12, atol=1e-12 * np.abs(e).max())
        assert np.isclose(w.pot[ip], pot, rtol=1e-12)


def test_tiles_partition_and_maximal(orc):
    s = workloads.explosion(n=8000, seed=8)
    d = orc.domain_decompose(s.x, s.y, s.z, s.q)
    for G in (1, 16, 64):
        tf, tc = orc.walk_tiles(d.keys, G)
        assert np.all(tc <= G) and np.all(tc >= 1)
        assert tf[0] == 0 and np.all(tf[1:] == tf[:-1] + tc[:-1]) and tf[-1] + tc[-1] == 8000
        # each tile is a whole key cell; its parent cell holds more than G
        # the packed rule merges runs of whole sibling tiles: every packed group
        # is a union of consecutive reference tiles and holds at most 64
        pf, pc = orc.walk_tiles(d.keys, G, pack=64)
        assert np.all(pc <= 64) and pf[0] == 0 and np.all(pf[1:] == pf[:-1] + pc[:-1])
        assert set(pf.tolist()) <= set(tf.tolist()) and len(pf) <= len(tf)
        for a, c in zip(tf[::53], tc[::53]):
            k = d.keys[a:a + c]
            for l in range(20, -1, -1):
                sh = np.uint64(3 * (20 - l))
                if np.all((k >> sh) == (k[0] >> sh)):
                    lvl = l
                    break
            for l in range(lvl, -1, -1):   # tile level: coarsest cell with <= G particles
                sh = np.uint64(3 * (20 - l))
                cell = d.keys >> sh
                cnt = int(np.sum(cell == (k[0] >> sh)))
                if cnt == c:
                    tl = l
                else:
                    assert cnt > G
                    break


def test_tree_accuracy_vs_direct(orc):
    s = workloads.explosion(n=20000, seed=12)
    d, t, w = orc.solve(s.x, s.y, s.z, s.q, theta2=0.36, eps2=s.eps2, tile_size=16)
    idx = np.arange(0, 20000, 97)
    e, pot = orc.direct(d.x, d.y, d.z, d.q, idx, s.eps2)
    rel = np.linalg.norm(w.e[idx] - e, axis=1) / np.linalg.norm(e, axis=1)
    assert np.median(rel) < 1e-3
    assert np.all(w.work == 1 + w.p2p + w.m2p)
    assert w.num_interactions == int((w.p2p + w.m2p).sum())
