"""Pins of the oracle's MAC-evaluation counter against the reference's rule
(module_tree_walk_simple.f90:343-371): a tile's visit of a twig counts one
evaluation per target, in tile order, up to and including the first target
that fails the MAC (the node is then opened for the whole tile), or the
tile's size when every target accepts it.  Leaves count nothing.  The GPU's
per-group counts are compared bit for bit with these.

- theta = 0 never accepts (theta2 d^2 > b^2 is false), so every twig of the
  tree is visited by every tile and fails at its first target: each tile
  counts exactly the tree's twig count;
- a root that every target accepts: four equal charges on the corners of a
  regular tetrahedron inscribed in the unit cube all lie at distance^2 0.75
  from the root's centre of charge; with theta^2 = 2, 2 * 0.75 exceeds
  boxlength2(0) = (1 + 2e-4)^2, so the one tile counts 4 (its size) and each
  target gets one multipole interaction and nothing else (at theta^2 = 1.3
  the root fails at the first target instead: count 1);
- the k-th target fails: one heavy charge and three light ones on four
  corners of the cube.  The root's centre of charge sits (nearly) on the
  heavy particle, the light ones lie 0.9 or more away and accept the root
  at theta^2 = 1.5, the heavy one fails (distance 0).  With the heavy
  particle at key rank k the tile counts k; the root's children are leaves,
  so nothing else is counted."""
import numpy as np
import pytest


def _solve(orc, pts, q, theta2, G):
    d = orc.domain_decompose(pts[:, 0], pts[:, 1], pts[:, 2], q)
    t = orc.tree_build(d.keys, d.x, d.y, d.z, d.q, d.box)
    w = orc.walk_group(d.x, d.y, d.z, t, theta2, 1e-6, tile_size=G, keys=d.keys)
    return d, t, w


@pytest.mark.parametrize("n,G", [(500, 16), (2000, 16), (777, 1), (3000, 64)])
def test_theta_zero_counts_every_twig_once_per_tile(orc, n, G):
    rng = np.random.default_rng(n + G)
    pts = rng.random((n, 3))
    q = rng.choice([-1.0, 1.0], n)
    d, t, w = _solve(orc, pts, q, 0.0, G)
    assert (w.tile_mac == t.ntwig).all()
    assert w.num_mac == len(w.tile_first) * t.ntwig
    assert (w.p2p == n).all() and (w.m2p == 0).all()


def test_root_accepted_by_every_target_counts_tile_size(orc):
    pts = np.array([[0.0, 0.0, 0.0], [1.0, 1.0, 0.0], [1.0, 0.0, 1.0], [0.0, 1.0, 1.0]])
    d, t, w = _solve(orc, pts, np.ones(4), 2.0, 4)
    assert len(w.tile_first) == 1 and w.tile_count[0] == 4
    assert w.tile_mac[0] == 4 and w.num_mac == 4
    assert (w.m2p == 1).all() and (w.p2p == 0).all()
    assert (w.work == 2.0).all()
    # just below the threshold the root opens at the first target
    d, t, w = _solve(orc, pts, np.ones(4), 1.3, 4)
    assert w.tile_mac[0] == 1 and (w.p2p == 4).all()


CORNERS = np.array([[0.05, 0.05, 0.05], [0.95, 0.05, 0.05], [0.05, 0.95, 0.05],
                    [0.95, 0.95, 0.95]])


@pytest.mark.parametrize("heavy", range(4))
def test_count_ends_at_first_failing_target(orc, heavy):
    q = np.full(4, 1e-9)
    q[heavy] = 1.0
    d, t, w = _solve(orc, CORNERS, q, 1.5, 4)
    k = int(np.nonzero(d.perm == heavy)[0][0]) + 1        # key rank of the heavy one, 1-based
    assert t.ntwig == 1                                     # the root; its children are leaves
    assert len(w.tile_first) == 1 and w.tile_count[0] == 4
    assert w.tile_mac[0] == k and w.num_mac == k
    # the root failed, so every target meets the four leaves
    assert (w.p2p == 4).all() and (w.m2p == 0).all()


def test_count_with_every_rank_seen(orc):
    # the four placements of the heavy particle put it at four different key
    # ranks, so counts 1..4 all occur
    ranks = set()
    for heavy in range(4):
        q = np.full(4, 1e-9)
        q[heavy] = 1.0
        d, t, w = _solve(orc, CORNERS, q, 1.5, 4)
        ranks.add(int(w.tile_mac[0]))
    assert ranks == {1, 2, 3, 4}
