"""Pins of the oracle's frozen-list evaluation and adjoint (the checker of the
GPU adjoint).

* the list evaluation reproduces the group walk bit for bit, and the
  fixed-topology moments reproduce the tree build bit for bit;
* one interaction: the adjoint's derivatives equal central differences of the
  reference M2P/P2P formulas (or_m2p / or_p2p, module_coulomb_kernels.f90);
* whole walk: the gradient equals fourth-order central differences
  (h = 1e-5) of the frozen-list evaluation to 1e-6;
* theta = 0: the gradient equals the closed-form gradient of the pair sum.
"""
import numpy as np
import pytest


def _setup(orc, n=2000, seed=3, theta2=0.36, eps2=1e-4, G=16):
    r = np.random.default_rng(seed)
    p = r.random((n, 3))
    q = r.standard_normal(n)
    d = orc.domain_decompose(p[:, 0], p[:, 1], p[:, 2], q)
    t = orc.tree_build(d.keys, d.x, d.y, d.z, d.q, d.box)
    tiles = orc.walk_tiles(d.keys, G)
    lists = orc.walk_lists(d.x, d.y, d.z, t, theta2, tiles)
    return r, d, t, tiles, lists


def test_lists_reproduce_walk(orc):
    r, d, t, tiles, lists = _setup(orc)
    w = orc.walk_group(d.x, d.y, d.z, t, 0.36, 1e-4, tiles=tiles)
    mom = orc.tree_moments(t, d.x, d.y, d.z, d.q)
    assert np.array_equal(mom[:, :15], t.mom[:, :15])
    e, pot = orc.eval_lists(d.x, d.y, d.z, t, mom, lists, 1e-4)
    assert np.array_equal(e, w.e) and np.array_equal(pot, w.pot)
    # one list entry per (tile, node): interactions = sum over tiles of size * entries
    per_tile = np.diff(lists.off)
    assert int(np.sum(per_tile * lists.tile_count)) == w.num_interactions


def _m2p_terms(orc, mom, d, eps2):
    e, phi = orc.m2p(mom, d, float(d @ d) + eps2)
    return e, phi


def test_term_adjoint_matches_kernel_differences(orc):
    r = np.random.default_rng(11)
    eps2 = 1e-3
    h = 1e-6
    for trial in range(20):
        mom = np.zeros(16)
        mom[orc.CHARGE] = r.standard_normal()
        mom[orc.DIPX:orc.DIPZ + 1] = 0.3 * r.standard_normal(3)
        mom[orc.QUADX:orc.ZXQUAD + 1] = 0.2 * r.standard_normal(6)
        d = r.standard_normal(3)
        d *= (0.5 + r.random()) / np.linalg.norm(d)
        gphi = r.standard_normal()
        u = r.standard_normal(3)
        p = mom[orc.DIPX:orc.DIPZ + 1]
        qt = mom[orc.QUADX:orc.ZXQUAD + 1]
        D, gq, gp, M = orc.adjoint_term(d, d @ d + eps2, gphi, u, mom[orc.CHARGE], p, qt)

        def L(mm, dd):
            e, phi = _m2p_terms(orc, mm, dd, eps2)
            return gphi * phi + u @ e

        def fd(fun):
            return (8 * (fun(h) - fun(-h)) - (fun(2 * h) - fun(-2 * h))) / (12 * h)

        for k in range(3):
            e_k = np.eye(3)[k]
            assert np.isclose(fd(lambda s: L(mom, d + s * e_k)), D[k], rtol=1e-7, atol=1e-9)
        slots = [orc.CHARGE, orc.DIPX, orc.DIPY, orc.DIPZ, orc.QUADX, orc.QUADY, orc.QUADZ,
                 orc.XYQUAD, orc.YZQUAD, orc.ZXQUAD]
        # dL/dqx = M_xx; dL/dxy = M_xy (Qt_xy = Qt_yx = xy/2)
        want = [gq, gp[0], gp[1], gp[2], M[0], M[1], M[2], M[3], M[4], M[5]]
        for slot, w in zip(slots, want):
            def f(s, slot=slot):
                mm = mom.copy()
                mm[slot] += s
                return L(mm, d)
            assert np.isclose(fd(f), w, rtol=1e-7, atol=1e-9), (trial, slot)
    # P2P: the M2P adjoint with zero dipole and quadrupole
    d = np.array([0.3, -0.2, 0.5])
    e, phi = orc.p2p(1.7, d, d @ d + eps2)
    D, gq, gp, M = orc.adjoint_term(d, d @ d + eps2, 1.0, np.zeros(3), 1.7)
    assert np.allclose(D, -e, rtol=1e-13)           # d(phi)/dd = -E
    assert np.isclose(gq, phi / 1.7, rtol=1e-13)


def test_adjoint_matches_frozen_finite_differences(orc):
    r, d, t, tiles, lists = _setup(orc)
    n = len(d.x)
    eps2 = 1e-4
    ge = r.standard_normal((n, 3))
    gpot = r.standard_normal(n)
    G = orc.adjoint_lists(d.x, d.y, d.z, d.q, t, lists, eps2, ge, gpot)
    X = [d.x.copy(), d.y.copy(), d.z.copy(), d.q.copy()]

    def loss(Xs):
        mom = orc.tree_moments(t, *Xs)
        e, pot = orc.eval_lists(Xs[0], Xs[1], Xs[2], t, mom, lists, eps2)
        return np.sum(e * ge) + np.sum(pot * gpot)

    h = 1e-5
    worst = 0.0
    for k in range(120):
        c, j = k % 4, int(r.integers(n))

        def f(s):
            Xs = [a.copy() for a in X]
            Xs[c][j] += s
            return loss(Xs)

        fd = (8 * (f(h) - f(-h)) - (f(2 * h) - f(-2 * h))) / (12 * h)
        scale = max(abs(G[c][j]), 1e-3 * np.sqrt(np.mean(G[c] ** 2)))
        worst = max(worst, abs(fd - G[c][j]) / scale)
    print(f"worst relative FD deviation {worst:.2e}")
    assert worst <= 1e-6


def test_theta0_adjoint_equals_pair_sum_gradient(orc):
    r = np.random.default_rng(13)
    n, eps2 = 300, 1e-3
    p = r.random((n, 3))
    q = r.standard_normal(n)
    d = orc.domain_decompose(p[:, 0], p[:, 1], p[:, 2], q)
    t = orc.tree_build(d.keys, d.x, d.y, d.z, d.q, d.box)
    tiles = orc.walk_tiles(d.keys, 16)
    lists = orc.walk_lists(d.x, d.y, d.z, t, 0.0, tiles)
    ge = r.standard_normal((n, 3))
    gpot = r.standard_normal(n)
    gx, gy, gz, gq = orc.adjoint_lists(d.x, d.y, d.z, d.q, t, lists, eps2, ge, gpot, sign=-1.0)
    # closed form, L = sum_i gpot_i phi_i + ge_i . E_i with phi, E of the pair
    # sum times the sign -1
    X = np.stack([d.x, d.y, d.z], 1)
    D = X[:, None, :] - X[None, :, :]                   # d_ij = x_i - x_j
    r2 = np.sum(D * D, 2) + eps2
    np.fill_diagonal(r2, 1.0)
    ri = 1.0 / np.sqrt(r2)
    np.fill_diagonal(ri, 0.0)
    r3, r5 = ri ** 3, ri ** 5
    gphi = -gpot
    u = -ge
    ud = np.einsum("ik,ijk->ij", u, D)
    qj = d.q[None, :]
    # dL/dd_ij for the (i, j) term
    T = qj[:, :, None] * (-gphi[:, None, None] * D * r3[:, :, None]
                          + u[:, None, :] * r3[:, :, None]
                          - 3.0 * ud[:, :, None] * D * r5[:, :, None])
    want_x = T.sum(1) - T.sum(0)
    want_q = (gphi[:, None] * ri + ud * r3).sum(0)
    got_x = np.stack([gx, gy, gz], 1)
    assert np.allclose(got_x, want_x, rtol=1e-10, atol=1e-10 * np.abs(want_x).max())
    assert np.allclose(gq, want_q, rtol=1e-10, atol=1e-10 * np.abs(want_q).max())
