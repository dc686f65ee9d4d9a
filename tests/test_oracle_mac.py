"""Pins of the oracle's MAC selection (coulomb/module_interaction_specific.f90:
256-275): 0 Barnes-Hut theta2 dist2 > boxlength2(level); 1 Bmax theta2 dist2 >
min(bmax^2, 3 boxlength2); any other value never accepts (the N^2 default).

- never-accept: every target meets every particle through its leaf, at any
  theta, and the fields are the direct pair sum;
- Bmax: a hand-built five-particle tree in which one twig holds a compact
  pair (bmax 0.05) in a cell of side 1/4.  For the target 0.3 away, theta^2
  d^2 = 0.0324 is below boxlength2 = 0.0625 (Barnes-Hut opens the twig: five
  leaves) and above bmax^2 = 0.0025 (Bmax accepts it: three leaves and one
  multipole)."""
import numpy as np
import pytest


def _sorted_index(d, i):
    return int(np.nonzero(d.perm == i)[0][0])


def test_never_accept_is_direct_sum(orc):
    rng = np.random.default_rng(5)
    n = 300
    x, y, z = rng.random((3, n))
    q = rng.uniform(-1.0, 1.0, n)
    d = orc.domain_decompose(x, y, z, q)
    t = orc.tree_build(d.keys, d.x, d.y, d.z, d.q, d.box)
    with orc.mac_select(2):
        w = orc.walk_group(d.x, d.y, d.z, t, 0.36, 1e-4, tile_size=16, keys=d.keys)
    assert (w.p2p == n).all() and (w.m2p == 0).all()
    e, pot = orc.direct(d.x, d.y, d.z, d.q, np.arange(n), 1e-4)
    assert np.abs(w.e - e).max() <= 1e-12 * np.abs(e).max()
    assert np.abs(w.pot - pot).max() <= 1e-12 * np.abs(pot).max()
    # the setting is scoped: Barnes-Hut again outside the block
    w0 = orc.walk_group(d.x, d.y, d.z, t, 0.36, 1e-4, tile_size=16, keys=d.keys)
    assert w0.m2p.sum() > 0


def test_bmax_accepts_compact_twig_that_barnes_hut_opens(orc):
    pts = np.array([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0],          # box corners
                    [0.3, 0.3, 0.3], [0.3, 0.3, 0.4],          # the pair, cell [1/4, 1/2)^3
                    [0.6, 0.3, 0.35]])                         # the target
    q = np.ones(5)
    d = orc.domain_decompose(pts[:, 0], pts[:, 1], pts[:, 2], q)
    t = orc.tree_build(d.keys, d.x, d.y, d.z, d.q, d.box)
    it = _sorted_index(d, 4)
    w0 = orc.walk_group(d.x, d.y, d.z, t, 0.36, 0.0, tile_size=1, keys=d.keys)
    with orc.mac_select(1):
        w1 = orc.walk_group(d.x, d.y, d.z, t, 0.36, 0.0, tile_size=1, keys=d.keys)
    assert (w0.p2p[it], w0.m2p[it]) == (5, 0)
    assert (w1.p2p[it], w1.m2p[it]) == (3, 1)
    # the multipole of the pair stands in for its two charges: the field at
    # the target differs from the Barnes-Hut (exact pair) result only by the
    # octupole remainder, (0.05 / 0.3)^3 of the pair's field
    pair = np.linalg.norm(pts[4] - pts[2]) ** -2
    assert np.linalg.norm(w1.e[it] - w0.e[it]) < 5e-3 * pair
    assert np.linalg.norm(w1.e[it] - w0.e[it]) > 0.0
