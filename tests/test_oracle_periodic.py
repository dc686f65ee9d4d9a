"""Pins of the oracle's periodic-boundary restatement (CPU).

- neighbour boxes: the order of module_mirror_boxes.f90:119-149 written out
  by hand for small cases;
- lattice coefficients: the reference's own pretabulated unit-cube table
  (tests/golden/lattice_unit_cube.json, written by make_lattice_fixture.py
  from module_fmm_framework.f90:260-508) and the scaling law of a cube of
  side a, ML(a) = ML(1) / a^(l+1);
- near field + lattice field: an Ewald sum (a different method for the same
  lattice sum) and the rock-salt Madelung constant;
- mirror-box walk: one box equals the plain group walk bit for bit; theta = 0
  over 27 boxes equals the 3x3x3 image sum.
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import periodic as P

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def table():
    with open(os.path.join(HERE, "golden", "lattice_unit_cube.json")) as f:
        return P.unit_cube_table(json.load(f))


@pytest.fixture(scope="module")
def ml_unit():
    return P.lattice_coefficients()


def test_neighbour_box_order():
    assert P.neighbour_boxes((1, 0, 0)).tolist() == [[-1, 0, 0], [1, 0, 0], [0, 0, 0]]
    assert P.neighbour_boxes((0, 0, 1)).tolist() == [[0, 0, -1], [0, 0, 1], [0, 0, 0]]
    b = P.neighbour_boxes((1, 1, 1))
    assert len(b) == 27
    assert b[:6].tolist() == [[-1, -1, -1], [1, 1, 1], [-1, -1, 0], [1, 1, 0], [-1, -1, 1], [1, 1, -1]]
    assert b[18:].tolist() == [[0, -1, -1], [0, 1, 1], [0, -1, 0], [0, 1, 0], [0, -1, 1], [0, 1, -1],
                               [0, 0, -1], [0, 0, 1], [0, 0, 0]]
    assert sorted(map(tuple, b.tolist())) == sorted(
        (i, j, k) for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1))
    assert (b[0:26:2] == -b[1:26:2]).all()
    assert len(P.neighbour_boxes((0, 1, 1), layers=2)) == 25


def _row_scale(tab):
    return {l: max(abs(tab[P.idx(l, m)]) for m in range(l + 1)) for l in range(P.LMAX_T + 1)}


def test_coefficients_match_reference_table(table, ml_unit):
    """Computed with multipole order 20 against the reference's table (order
    50): rows l >= 4 agree to 1e-8; l = 0 (the divergent monopole sum, grown
    by 9 per iteration) to 1e-13; l = 2, 3 vanish by cubic symmetry.  Row
    l = 1 is excluded: the iteration multiplies it by 3 per step, so it holds
    rounding noise times 3^32 (it only meets the zero monopole and dipole of
    the central box)."""
    assert abs(ml_unit[0] - table[0]) <= 1e-13 * abs(table[0])
    for l in (2, 3):
        for m in range(l + 1):
            assert abs(ml_unit[P.idx(l, m)]) < 1e-12
    rows = _row_scale(table)
    worst = 0.0
    for l in range(4, P.LMAX_T + 1):
        scale = max(rows[k] for k in range(l - 2, min(l + 2, P.LMAX_T) + 1))
        for m in range(l + 1):
            i = P.idx(l, m)
            ref = table[i]
            err = abs(ml_unit[i] - ref) / (abs(ref) if ref != 0 else scale)
            worst = max(worst, err)
    assert worst < 1e-8, worst


def test_coefficients_scaling_law(table, ml_unit):
    """Entries that cubic symmetry leaves nonzero (those of the table)."""
    a = 1.7
    ml_a = P.lattice_coefficients(lattice=(a, a, a))
    nz = 0
    for l in [0] + list(range(4, P.LMAX_T + 1)):
        for m in range(l + 1):
            i = P.idx(l, m)
            if table[i] != 0:
                want = ml_unit[i] / a ** (l + 1)
                assert abs(ml_a[i] - want) <= 1e-12 * abs(want), (l, m)
                nz += 1
    assert nz > 100


def _neutral(n, seed, box=1.0, origin=0.0):
    rng = np.random.default_rng(seed)
    x, y, z = origin + box * rng.random((3, n))
    q = np.where(np.arange(n) % 2 == 0, 1.0, -1.0) * rng.uniform(0.5, 1.5, n)
    q -= q.mean()
    return x, y, z, q


def _input_order(d, w):
    n = len(d.perm)
    e = np.empty((n, 3))
    p = np.empty(n)
    e[d.perm] = w.e
    p[d.perm] = w.pot
    return e, p


@pytest.mark.parametrize("box,origin", [(1.0, 0.0), (2.0, -1.0)])
def test_lattice_sum_equals_ewald(table, box, origin):
    """Mirror walks (theta = 0) + lattice field = tin-foil Ewald sum: the
    field to 1e-6 of its rms, the potential up to the constant
    -(2 pi / 3V) sum q |r - c|^2 (the extrinsic quadrupole term by which the
    intrinsic lattice sum differs from the tin-foil one), energies to 1e-7."""
    x, y, z, q = _neutral(48, 3, box, origin)
    ml = table if box == 1.0 else P.lattice_coefficients(lattice=(box, box, box))
    d, t, w = P.solve_periodic(x, y, z, q, 0.0, 0.0, ml, lattice=(box,) * 3, origin=(origin,) * 3)
    e, p = _input_order(d, w)
    e_ew, p_ew = P.ewald(x - origin, y - origin, z - origin, q, box=box)
    rms = math.sqrt((e_ew ** 2).mean())
    assert np.abs(e - e_ew).max() <= 1e-6 * rms
    c = origin + 0.5 * box
    r2 = (x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2
    shift = -2 * math.pi / (3 * box ** 3) * (q * r2).sum()
    dev = p - (p_ew + shift)
    assert np.abs(dev).max() <= 1e-5 * math.sqrt((p_ew ** 2).mean())
    u, u_ew = 0.5 * (q * p).sum(), 0.5 * (q * p_ew).sum()
    assert abs(u - u_ew) <= 1e-7 * abs(u_ew)


def _dipole_terms(x, y, z, q, box, origin):
    c = origin + 0.5 * box
    r = np.stack([x - c, y - c, z - c], 1)
    return r, (q[:, None] * r).sum(0), (q * (r * r).sum(1)).sum(), box ** 3


@pytest.mark.parametrize("box,origin", [(1.0, 0.0), (2.0, -1.0)])
def test_extrinsic_corrections_against_ewald(table, box, origin):
    """The three extrinsic-to-intrinsic corrections of the reference
    (module_fmm_framework.f90:42-46, 807-829) against the tin-foil Ewald sum,
    theta = 0, one mirror layer.  Redlack-Grindlay (E += 2 pf D, phi += -2 pf
    r.D + pf Q, pf = 2 pi / 3V) is the tin-foil sum itself: the field to 1e-7
    of its rms; the potential too in the unit cube, whose tabulated l = 1
    lattice row is zero, and up to a constant in a computed lattice (the
    reference's l = 1 row holds rounding noise, which multiplies the box
    dipole in the constant term).  No correction is the extrinsic sum, which
    differs from the tin-foil one by closed forms: E - (4 pi / 3V) D and
    phi + (4 pi / 3V) r.D - pf Q.  A dropped factor 2 or a sign in either
    correction fails these by O(1)."""
    x, y, z, q = _neutral(48, 3, box, origin)
    ml = table if box == 1.0 else P.lattice_coefficients(lattice=(box, box, box))
    e_ew, p_ew = P.ewald(x - origin, y - origin, z - origin, q, box=box)
    rms = math.sqrt((e_ew ** 2).mean())
    prms = math.sqrt((p_ew ** 2).mean())
    r, D, Q, V = _dipole_terms(x, y, z, q, box, origin)
    pf = 2 * math.pi / (3 * V)
    kw = dict(lattice=(box,) * 3, origin=(origin,) * 3)
    d, t, w = P.solve_periodic(x, y, z, q, 0.0, 0.0, ml, correction=P.EXTRINSIC_REDLACK, **kw)
    e, p = _input_order(d, w)
    assert np.abs(e - e_ew).max() <= 1e-7 * rms
    dev = p - p_ew
    if box == 1.0:
        assert np.abs(dev).max() <= 1e-7 * prms
    else:
        assert np.abs(dev - dev.mean()).max() <= 1e-7 * prms
    d, t, w = P.solve_periodic(x, y, z, q, 0.0, 0.0, ml, correction=P.EXTRINSIC_NONE, **kw)
    e, p = _input_order(d, w)
    assert np.abs(e - (e_ew - 2 * pf * D)).max() <= 1e-7 * rms
    dev = p - (p_ew + 2 * pf * (r @ D) - pf * Q)
    if box == 1.0:
        assert np.abs(dev).max() <= 1e-7 * prms
    else:
        assert np.abs(dev - dev.mean()).max() <= 1e-7 * prms


@pytest.mark.slow
def test_two_mirror_layers_against_ewald():
    """mirror_box_layers = 2 (ws = 2 in the far field, module_fmm_framework.
    f90:111, 994): 125 mirror walks at theta = 0 plus the lattice beyond
    (coefficients with the 5x-scaled L* terms, 124 x 125 per entry) equal the
    tin-foil Ewald sum, with the fictitious charges (field 1e-9 of its rms,
    potential up to -pf Q to 1e-9) and with Redlack-Grindlay (field 1e-9,
    potential up to a constant)."""
    ml2 = P.lattice_coefficients(layers=2)
    x, y, z, q = _neutral(48, 3)
    e_ew, p_ew = P.ewald(x, y, z, q)
    rms = math.sqrt((e_ew ** 2).mean())
    prms = math.sqrt((p_ew ** 2).mean())
    r, D, Q, V = _dipole_terms(x, y, z, q, 1.0, 0.0)
    pf = 2 * math.pi / (3 * V)
    d, t, w = P.solve_periodic(x, y, z, q, 0.0, 0.0, ml2, layers=2)
    e, p = _input_order(d, w)
    assert np.abs(e - e_ew).max() <= 1e-9 * rms
    assert np.abs(p - (p_ew - pf * Q)).max() <= 1e-9 * prms
    d, t, w = P.solve_periodic(x, y, z, q, 0.0, 0.0, ml2, layers=2, correction=P.EXTRINSIC_REDLACK)
    e, p = _input_order(d, w)
    assert np.abs(e - e_ew).max() <= 1e-9 * rms
    dev = p - p_ew
    assert np.abs(dev - dev.mean()).max() <= 1e-7 * prms


def test_rock_salt_madelung(table):
    x, y, z, q = P.rock_salt(1)
    x, y, z = x + 0.125, y + 0.125, z + 0.125      # no ion on a box corner
    d, t, w = P.solve_periodic(x, y, z, q, 0.0, 0.0, table)
    madelung = -w.pot * d.q * 0.5                   # nearest-neighbour distance 1/2
    assert np.abs(madelung - P.MADELUNG_NACL).max() < 2e-7


def test_walk_boxes_one_box_is_group_walk(orc):
    x, y, z, q = _neutral(3000, 5)
    d = orc.domain_decompose(x, y, z, q)
    t = orc.tree_build(d.keys, d.x, d.y, d.z, d.q, d.box)
    w0 = orc.walk_group(d.x, d.y, d.z, t, 0.36, 1e-4, keys=d.keys)
    w1 = P.walk_group_boxes(d.x, d.y, d.z, t, 0.36, 1e-4, np.zeros((1, 3)), keys=d.keys)
    assert np.array_equal(w0.e, w1.e) and np.array_equal(w0.pot, w1.pot)
    assert np.array_equal(w0.p2p, w1.p2p) and np.array_equal(w0.m2p, w1.m2p)
    assert np.array_equal(w0.tile_mac, w1.tile_mac)


def test_walk_boxes_theta0_is_image_sum(orc):
    x, y, z, q = _neutral(200, 7)
    d = orc.domain_decompose(x, y, z, q)
    t = orc.tree_build(d.keys, d.x, d.y, d.z, d.q, d.box)
    vb = P.lattice_vectors((1.0, 1.0, 1.0), P.neighbour_boxes())
    w = P.walk_group_boxes(d.x, d.y, d.z, t, 0.0, 0.0, vb, keys=d.keys)
    pos = np.stack([d.x, d.y, d.z], 1)
    e = np.zeros_like(pos)
    p = np.zeros(len(q))
    for v in vb:
        dd = pos[:, None, :] - (pos[None, :, :] + v)
        r2 = (dd * dd).sum(2)
        if not v.any():
            np.fill_diagonal(r2, np.inf)
        rinv = 1.0 / np.sqrt(r2)
        p += (rinv * d.q[None, :]).sum(1)
        e += ((rinv ** 3 * d.q[None, :])[:, :, None] * dd).sum(1)
    assert np.abs(w.e - e).max() <= 1e-12 * np.abs(e).max()
    assert np.abs(w.pot - p).max() <= 1e-12 * np.abs(p).max()
    assert (w.p2p == 27 * len(q)).all() and (w.m2p == 0).all()
