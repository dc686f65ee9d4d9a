"""Pins the oracle's box and Hilbert-key restatement (module_box.f90:48-93,
module_spacefilling.f90:224-257, :380-440, :445-565) by invariants that a
wrong table entry, bit order or coordinate swap breaks."""
import numpy as np
import pytest

from pepc_amd import workloads


def test_box_margin_closed_form(orc):
    r = np.random.default_rng(0)
    x, y, z = r.random(1000) * 3 - 1, r.random(1000) * 0.5, r.random(1000) + 7
    b = orc.box_create(x, y, z)
    bmin, bmax, bsize = b.as_arrays()
    for d, c in enumerate((x, y, z)):
        s0 = c.max() - c.min()
        assert bmin[d] == c.min() - s0 / 10000.0
        assert bmax[d] == c.max() + s0 / 10000.0
        assert bsize[d] == bmax[d] - bmin[d]
    bc = orc.box_create(x, y, z, force_cubic=True)
    assert np.all(np.array(bc.boxmin[:]) == bc.boxmin[0])


def test_hilbert_level3_bijective_and_face_adjacent(orc):
    # all 8^3 cells of level 3: prefixes form a permutation and consecutive
    # cells along the curve share a face (the defining Hilbert property)
    cells = {}
    for i in range(8):
        for j in range(8):
            for k in range(8):
                key = orc.hilbert_key(i << 17, j << 17, k << 17)
                assert orc.level_from_key(key) == 20
                cells[key >> 51] = (i, j, k)
    assert sorted(cells) == list(range(8 ** 3, 2 * 8 ** 3))
    order = [cells[p] for p in sorted(cells)]
    for a, b in zip(order, order[1:]):
        assert sum(abs(u - v) for u, v in zip(a, b)) == 1, (a, b)


def test_hilbert_fine_adjacency_random_runs(orc):
    # at full resolution: walk a contiguous key range and check adjacency
    r = np.random.default_rng(1)
    for _ in range(20):
        ic = r.integers(0, 1 << 20, 3)
        k0 = orc.hilbert_key(*ic)
        prev = orc.key_to_intcoord(k0)
        assert np.array_equal(prev, ic)
        for dk in range(1, 40):
            k = k0 + dk
            if k >= (1 << 61):
                break
            cur = orc.key_to_intcoord(k)
            assert np.abs(cur - prev).sum() == 1
            assert orc.hilbert_key(*cur) == k
            prev = cur


def test_key_roundtrip_on_particles(orc):
    s = workloads.explosion(n=20000, seed=11)
    b = orc.box_create(s.x, s.y, s.z)
    keys = orc.compute_keys(s.x, s.y, s.z, b)
    assert np.all((keys >> np.uint64(60)) == 1)
    bmin, _, bsize = b.as_arrays()
    for i in range(0, 20000, 997):
        c = orc.key_to_coord(b, keys[i])
        # the cell centre lies within half a cell of the particle, per axis
        cell = bsize / 2.0 ** 20
        p = np.array([s.x[i], s.y[i], s.z[i]])
        assert np.all(np.abs(c - p) <= 0.5 * cell * (1 + 1e-9))
        # and maps back to the same key
        k2 = orc.compute_keys(np.array([c[0]]), np.array([c[1]]), np.array([c[2]]), b)[0]
        assert k2 == keys[i]


def test_level1_octant_order(orc):
    # level-1 digit = position of the octant along the curve, from the
    # reference table CI3 = [0,1,3,2,7,6,4,5] indexed by cval = z<<2|y<<1|x
    ci3 = [0, 1, 3, 2, 7, 6, 4, 5]
    for cval in range(8):
        ix, iy, iz = (cval & 1), (cval >> 1) & 1, (cval >> 2) & 1
        key = orc.hilbert_key(ix << 19, iy << 19, iz << 19)
        assert (key >> 57) & 7 == ci3[cval]
