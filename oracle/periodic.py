"""Periodic boundaries: ctypes front end of the oracle's restatement
(pepc_oracle_periodic.c) and an independent Ewald sum.

TEST INFRASTRUCTURE ONLY: the checker of the GPU periodic path; the product
package never imports it.

The restatement follows src/treecode/lpepcsrc/module_mirror_boxes.f90 (mirror
boxes) and src/interaction_specific/coulomb/module_fmm_framework.f90 (lattice
contribution, fictitious-charge correction).  The Ewald sum below is a
different method for the same conditionally convergent lattice sum, with the
conducting (tin-foil) boundary; it pins the restatement together with the
rock-salt Madelung constant.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import oracle as _o

LMAX_M = 20
LMAX_T = 40
NM = 231
NT = 861
MAXITER = 32

_f64 = _o._f64
_i32 = _o._i32
_bound = False


def _lib():
    global _bound
    L = _o.lib()
    if not _bound:
        L.or_neighbour_boxes.argtypes = [_i32, C.c_int32, C.c_void_p]
        L.or_neighbour_boxes.restype = C.c_int
        L.or_lattice_coefficients.argtypes = [_f64, _i32, C.c_int32, C.c_int32, _f64]
        L.or_lattice_coefficients.restype = C.c_int
        L.or_lattice_timestep.argtypes = [_f64, _f64, _i32, _f64, C.c_int64, _f64, _f64, _f64,
                                          _f64, C.c_int32, _f64, _f64, _f64, _i32, _f64, _f64]
        L.or_lattice_field.argtypes = [_f64, _f64, _i32, C.c_int32, _f64, C.c_int32, _f64,
                                       C.c_int32, _f64, C.c_double, C.c_int64, _f64, _f64, _f64,
                                       _f64, _f64, _f64, _f64]
        L.or_walk_group_boxes.argtypes = [C.c_int64, _f64, _f64, _f64, C.c_int64, _i32, _o._i64,
                                          _o._i64, _o._i64, _f64, _f64, C.c_double, C.c_double,
                                          C.c_double, C.c_int64, _o._i64, _o._i64, C.c_int64,
                                          C.c_int, C.c_int64, _f64, _f64, _f64, _f64, _f64,
                                          _f64, _i32, _i32, _o._i64, _o._i64]
        L.or_walk_group_boxes.restype = C.c_int
        _bound = True
    return L


def idx(l, m):
    return l * (l + 1) // 2 + m


def _lat(lattice):
    a = np.asarray(lattice, np.float64)
    if a.shape == (3,):
        a = np.diag(a)
    return np.ascontiguousarray(a.reshape(9))


def _per(periodicity):
    return np.ascontiguousarray(np.asarray(periodicity, np.int32).reshape(3))


def neighbour_boxes(periodicity=(1, 1, 1), layers=1) -> np.ndarray:
    """(nb, 3) integer lattice indices, module_mirror_boxes.f90:104-151 order."""
    per = _per(periodicity)
    nb = _lib().or_neighbour_boxes(per, layers, None)
    out = np.zeros(3 * nb, np.int32)
    _lib().or_neighbour_boxes(per, layers, out.ctypes.data)
    return out.reshape(nb, 3)


def lattice_vectors(lattice, boxes) -> np.ndarray:
    lat = _lat(lattice).reshape(3, 3)
    b = np.asarray(boxes, np.float64)
    return b[:, 0:1] * lat[0] + b[:, 1:2] * lat[1] + b[:, 2:3] * lat[2]


def lattice_coefficients(lattice=(1.0, 1.0, 1.0), periodicity=(1, 1, 1), layers=1,
                         maxiter=MAXITER) -> np.ndarray:
    """calc_lattice_coefficients: complex (NT,) Taylor table."""
    out = np.zeros(2 * NT)
    rc = _lib().or_lattice_coefficients(_lat(lattice), _per(periodicity), layers, maxiter, out)
    if rc:
        raise MemoryError("or_lattice_coefficients")
    return out.view(np.complex128)


def unit_cube_table(fixture) -> np.ndarray:
    """The reference's pretabulated unit-cube coefficients (fixture entries
    [l, m, value]) as a complex (NT,) table; orders above LMAX_T dropped."""
    ml = np.zeros(NT, np.complex128)
    for l, m, v in fixture["entries"]:
        if l <= LMAX_T:
            ml[idx(l, m)] = v
    return ml


# extrinsic-to-intrinsic corrections (module_fmm_framework.f90:42-46)
EXTRINSIC_NONE, EXTRINSIC_REDLACK, EXTRINSIC_FICTCHARGE = 0, 1, 2


@dataclass
class Step:
    omega: np.ndarray      # complex (NM,)
    mu: np.ndarray         # complex (NT,)
    fict: np.ndarray       # (nfict, 4) x, y, z, q
    dipole: np.ndarray     # (3,)
    quad_trace: float = 0.0


class Lattice:
    """3-D periodic far field with an extrinsic-to-intrinsic correction
    (default: fictitious charges, the reference's default)."""

    def __init__(self, ml, lattice=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0),
                 periodicity=(1, 1, 1), layers=1, correction=EXTRINSIC_FICTCHARGE):
        self.ml = np.ascontiguousarray(np.asarray(ml, np.complex128))
        self.lat = _lat(lattice)
        self.origin = np.ascontiguousarray(np.asarray(origin, np.float64))
        self.per = _per(periodicity)
        self.layers = int(layers)
        self.correction = int(correction)

    def timestep(self, x, y, z, q) -> Step:
        om = np.zeros(2 * NM)
        mu = np.zeros(2 * NT)
        fict = np.zeros(16)
        nf = np.zeros(1, np.int32)
        dip = np.zeros(3)
        qt = np.zeros(1)
        c = _o._c
        _lib().or_lattice_timestep(self.lat, self.origin, self.per, self.ml.view(np.float64),
                                   len(q), c(x, np.float64), c(y, np.float64), c(z, np.float64),
                                   c(q, np.float64), self.correction, om, mu, fict, nf, dip, qt)
        self.step = Step(om.view(np.complex128), mu.view(np.complex128),
                         fict[:4 * nf[0]].reshape(-1, 4).copy(), dip, float(qt[0]))
        return self.step

    def field(self, x, y, z):
        """(e (n,3), pot (n,)) of the lattice contribution at positions."""
        n = len(x)
        ex, ey, ez, pot = (np.zeros(n) for _ in range(4))
        st = self.step
        fict = np.zeros(16)
        fict[:st.fict.size] = st.fict.ravel()
        c = _o._c
        _lib().or_lattice_field(self.lat, self.origin, self.per, self.layers,
                                np.ascontiguousarray(st.mu).view(np.float64), self.correction,
                                fict, len(st.fict), np.ascontiguousarray(st.dipole, np.float64),
                                float(st.quad_trace), n, c(x, np.float64), c(y, np.float64),
                                c(z, np.float64), ex, ey, ez, pot)
        return np.stack([ex, ey, ez], 1), pot


def walk_group_boxes(x, y, z, tree, theta2, eps2, vbox, tile_size=16, sign=1.0, nthreads=0,
                     keys=None, tiles=None, pack=0):
    """Group walk over every box of vbox (nb, 3) (module_libpepc_main.f90:184-187)."""
    n = len(x)
    if tiles is None:
        tiles = _o.walk_tiles(keys, tile_size, pack)
    tf, tc = tiles
    ex, ey, ez, pot, work = (np.zeros(n) for _ in range(5))
    p2p = np.zeros(n, np.int32)
    m2p = np.zeros(n, np.int32)
    tmac = np.zeros(len(tf), np.int64)
    tot = np.zeros(3, np.int64)
    vb = np.ascontiguousarray(np.asarray(vbox, np.float64).reshape(-1))
    c = _o._c
    rc = _lib().or_walk_group_boxes(n, c(x, np.float64), c(y, np.float64), c(z, np.float64),
                                    tree.nnodes, tree.level, tree.leaves, tree.first_child,
                                    tree.next_sibling, np.ascontiguousarray(tree.mom.ravel()),
                                    tree.b2, float(theta2), float(eps2), float(sign), len(tf), tf,
                                    tc, 1, int(nthreads), len(vb) // 3, vb, ex, ey, ez, pot, work,
                                    p2p, m2p, tmac, tot)
    if rc:
        raise ValueError("walk did not reach every leaf")
    return _o.WalkResult(np.stack([ex, ey, ez], 1), pot, work, p2p, m2p, tf, tc, tmac,
                         int(tot[0]), int(tot[1]), int(tot[2]))


def walk_sinks(sx, sy, sz, tree, nsources, theta2, eps2, vbox=((0.0, 0.0, 0.0),), sign=1.0,
               nthreads=0, tiles=None):
    """Walk of sink points through a tree of nsources particles: every sink
    its own tile (module_tree_walk_simple.f90:155-159) unless tiles are
    given; self = zero separation from a leaf's particle (:331-340)."""
    m = len(sx)
    if tiles is None:
        tiles = (np.arange(m, dtype=np.int64), np.ones(m, dtype=np.int64))
    tf, tc = tiles
    ex, ey, ez, pot, work = (np.zeros(m) for _ in range(5))
    p2p = np.zeros(m, np.int32)
    m2p = np.zeros(m, np.int32)
    tmac = np.zeros(len(tf), np.int64)
    tot = np.zeros(3, np.int64)
    vb = np.ascontiguousarray(np.asarray(vbox, np.float64).reshape(-1))
    c = _o._c
    rc = _lib().or_walk_group_boxes(nsources, c(sx, np.float64), c(sy, np.float64),
                                    c(sz, np.float64), tree.nnodes, tree.level, tree.leaves,
                                    tree.first_child, tree.next_sibling,
                                    np.ascontiguousarray(tree.mom.ravel()), tree.b2, float(theta2),
                                    float(eps2), float(sign), len(tf), c(tf, np.int64),
                                    c(tc, np.int64), 1, int(nthreads), len(vb) // 3, vb, ex, ey, ez,
                                    pot, work, p2p, m2p, tmac, tot)
    if rc:
        raise ValueError("walk did not reach every leaf")
    return _o.WalkResult(np.stack([ex, ey, ez], 1), pot, work, p2p, m2p, tf, tc, tmac,
                         int(tot[0]), int(tot[1]), int(tot[2]))


def solve_periodic(x, y, z, q, theta2, eps2, ml, lattice=(1.0, 1.0, 1.0),
                   origin=(0.0, 0.0, 0.0), layers=1, far_field=True, tile_size=16, sign=1.0,
                   nthreads=0, correction=EXTRINSIC_FICTCHARGE):
    """Grow + traverse with 3-D periodicity (libpepc_traverse_tree,
    module_libpepc_main.f90:161-200): mirror-box walks, then the lattice
    contribution of every particle (calc_force_per_particle,
    coulomb/module_interaction_specific.f90:393-427).  Sorted order."""
    d = _o.domain_decompose(x, y, z, q)
    t = _o.tree_build(d.keys, d.x, d.y, d.z, d.q, d.box)
    boxes = neighbour_boxes((1, 1, 1), layers)
    vb = lattice_vectors(lattice, boxes)
    w = walk_group_boxes(d.x, d.y, d.z, t, theta2, eps2, vb, tile_size=tile_size, sign=sign,
                         nthreads=nthreads, keys=d.keys)
    if far_field:
        lat = Lattice(ml, lattice, origin, (1, 1, 1), layers, correction)
        lat.timestep(d.x, d.y, d.z, d.q)
        e, p = lat.field(d.x, d.y, d.z)
        w.e = w.e + sign * e
        w.pot = w.pot + sign * p
    return d, t, w


# ------------------------------------------------------------------ Ewald sum
def ewald(x, y, z, q, box=1.0, alpha=None, tol=1e-13):
    """Ewald sum for a neutral system in a cubic box of side `box`, tin-foil
    boundary: field and potential at every particle from all other particles
    and all periodic images (own image included, own Coulomb term excluded).
    """
    pos = np.stack([np.asarray(x, float), np.asarray(y, float), np.asarray(z, float)], 1)
    q = np.asarray(q, float)
    n = len(q)
    L = float(box)
    if alpha is None:
        alpha = 5.6 / L
    s = math.sqrt(-math.log(tol))
    nr = int(math.ceil(s / (alpha * L))) + 1
    kmax = int(math.ceil(s * alpha * L / math.pi)) + 1
    phi = np.zeros(n)
    e = np.zeros((n, 3))
    erfc = np.vectorize(math.erfc)
    rng = range(-nr, nr + 1)
    for a in rng:
        for b in rng:
            for c in rng:
                v = np.array([a, b, c], float) * L
                d = pos[:, None, :] - (pos[None, :, :] + v)
                r = np.sqrt((d * d).sum(2))
                if a == b == c == 0:
                    np.fill_diagonal(r, np.inf)
                ok = ~np.isinf(r)
                rr = np.where(ok, r, 1.0)
                ec = np.where(ok, erfc(alpha * rr), 0.0)
                g = np.where(ok, ec / rr ** 3 + 2 * alpha / math.sqrt(math.pi)
                             * np.exp(-(alpha * rr) ** 2) / rr ** 2, 0.0)
                phi += (np.where(ok, ec / rr, 0.0) * q[None, :]).sum(1)
                e += ((g * q[None, :])[:, :, None] * d).sum(1)
    kv = 2 * math.pi / L
    ks = np.arange(-kmax, kmax + 1)
    K = np.stack(np.meshgrid(ks, ks, ks, indexing="ij"), -1).reshape(-1, 3)
    K = K[np.any(K != 0, 1)] * kv
    k2 = (K * K).sum(1)
    f = 4 * math.pi / L ** 3 * np.exp(-k2 / (4 * alpha * alpha)) / k2
    ph = pos @ K.T                                       # (n, nk)
    cs, sn = np.cos(ph), np.sin(ph)
    sc, ss = q @ cs, q @ sn                               # (nk,)
    phi += (cs * (f * sc) + sn * (f * ss)).sum(1)
    e += (sn * (f * sc) - cs * (f * ss)) @ K
    phi -= 2 * alpha / math.sqrt(math.pi) * q
    return e, phi


def rock_salt(cells=1, box=1.0):
    """Rock-salt ions in a cubic box: 2*cells ions per edge, alternating
    charges +-1, nearest-neighbour distance box / (2 cells)."""
    m = 2 * cells
    g = np.arange(m) * (box / m)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    i = np.arange(m)
    I, J, K = np.meshgrid(i, i, i, indexing="ij")
    q = np.where((I + J + K) % 2 == 0, 1.0, -1.0)
    return X.ravel(), Y.ravel(), Z.ravel(), q.ravel()


MADELUNG_NACL = 1.7475645946331821906362120355443974   # OEIS A085469
