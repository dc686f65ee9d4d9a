"""ctypes front end of the oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` as the checker.  The product package
(``pepc_amd``) never imports this module.

Each function mirrors one routine of the reference; see pepc_oracle.h for the
file:line each follows and for how each is pinned.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libref_sort.so")

MAXLEVEL = 20
MOM = 16
(COCX, COCY, COCZ, CHARGE, ABSCHARGE, DIPX, DIPY, DIPZ, QUADX, QUADY, QUADZ,
 XYQUAD, YZQUAD, ZXQUAD, BMAX) = range(15)

_P = np.ctypeslib.ndpointer
_f64 = _P(dtype=np.float64, flags="C_CONTIGUOUS")
_i64 = _P(dtype=np.int64, flags="C_CONTIGUOUS")
_u64 = _P(dtype=np.uint64, flags="C_CONTIGUOUS")
_i32 = _P(dtype=np.int32, flags="C_CONTIGUOUS")


class Box(C.Structure):
    _fields_ = [("boxmin", C.c_double * 3), ("boxmax", C.c_double * 3),
                ("boxsize", C.c_double * 3)]

    def as_arrays(self):
        return (np.array(self.boxmin[:]), np.array(self.boxmax[:]), np.array(self.boxsize[:]))


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle oracle)")
        L = C.CDLL(LIB_PATH)
        L.or_box_create.argtypes = [C.c_int64, _f64, _f64, _f64, C.c_int, C.POINTER(Box)]
        L.or_box_create.restype = C.c_int
        L.or_compute_keys.argtypes = [C.c_int64, _f64, _f64, _f64, C.POINTER(Box), _u64]
        L.or_intcoord_to_key_hilbert3d.argtypes = [C.c_int64] * 3
        L.or_intcoord_to_key_hilbert3d.restype = C.c_uint64
        L.or_key_to_intcoord_hilbert.argtypes = [C.c_uint64, _i64]
        L.or_key_to_coord.argtypes = [C.POINTER(Box), C.c_uint64, _f64]
        L.or_level_from_key.argtypes = [C.c_uint64]
        L.or_level_from_key.restype = C.c_int
        L.or_sort_radix_ip.argtypes = [C.c_int64, _u64, _i64, C.c_void_p]
        L.or_dupfix.argtypes = [C.c_int64, _u64, _f64, _f64, _f64, C.POINTER(Box)]
        L.or_dupfix.restype = C.c_int64
        L.or_domain_decompose.argtypes = [C.c_int64, _f64, _f64, _f64, _f64, C.c_int,
                                          C.POINTER(Box), _u64, _i64, _f64, _f64, _f64, _f64,
                                          C.POINTER(C.c_int64)]
        L.or_domain_decompose.restype = C.c_int
        L.or_tree_build.argtypes = [C.c_int64, _u64, _f64, _f64, _f64, _f64, C.c_int64,
                                    _u64, _i32, _i64, _i64, _i64, _i64, _i32, _f64, _i64, _i64]
        L.or_tree_build.restype = C.c_int64
        L.or_boxlength2.argtypes = [C.POINTER(Box), _f64]
        L.or_walk_tiles.argtypes = [C.c_int64, _u64, C.c_int64, C.c_int64, _i64, _i64]
        L.or_walk_tiles.restype = C.c_int64
        L.or_walk_group.argtypes = [C.c_int64, _f64, _f64, _f64, C.c_int64, _i32, _i64, _i64, _i64,
                                    _f64, _f64, C.c_double, C.c_double, C.c_double,
                                    C.c_int64, _i64, _i64, C.c_int64, C.c_int,
                                    _f64, _f64, _f64, _f64, _f64, _i32, _i32, _i64, _i64]
        L.or_walk_group.restype = C.c_int
        L.or_set_mac_select.argtypes = [C.c_int]
        L.or_set_mac_select.restype = None
        L.or_get_mac_select.argtypes = []
        L.or_get_mac_select.restype = C.c_int
        L.or_p2p.argtypes = [_f64, _f64, C.c_double, _f64, _f64]
        L.or_m2p.argtypes = [_f64, _f64, C.c_double, _f64, _f64]
        L.or_direct.argtypes = [C.c_int64, _f64, _f64, _f64, _f64, C.c_int64, _i64,
                                C.c_double, C.c_double, C.c_int, _f64, _f64, _f64, _f64]
        L.or_walk_lists.argtypes = [_f64, _f64, _f64, _i32, _i64, _i64, _f64, _f64, C.c_double,
                                    C.c_int64, _i64, _i64, C.c_int64, C.c_void_p, C.c_void_p]
        L.or_walk_lists.restype = C.c_int64
        L.or_tree_moments.argtypes = [C.c_int64, _i64, _i64, _i64, _f64, _f64, _f64, _f64, _f64]
        L.or_eval_lists.argtypes = [_f64, _f64, _f64, _i64, _i64, _f64, C.c_double, C.c_double,
                                    C.c_int64, _i64, _i64, _i64, _i64, _f64, _f64, _f64, _f64]
        L.or_adjoint_term.argtypes = [_f64, C.c_double, C.c_double, _f64, C.c_double, _f64, _f64,
                                      _f64, _f64, _f64, _f64]
        L.or_adjoint_lists.argtypes = [C.c_int64, _f64, _f64, _f64, _f64, C.c_int64, _i64, _i64,
                                       _i64, _f64, C.c_double, C.c_double, C.c_int64, _i64, _i64,
                                       _i64, _i64, _f64, _f64, _f64, _f64, _f64, _f64, _f64, _f64]
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def box_create(x, y, z, force_cubic=False) -> Box:
    b = Box()
    rc = lib().or_box_create(len(x), _c(x, np.float64), _c(y, np.float64), _c(z, np.float64),
                             int(force_cubic), C.byref(b))
    if rc:
        raise ValueError(f"or_box_create failed ({rc})")
    return b


def compute_keys(x, y, z, b: Box) -> np.ndarray:
    keys = np.empty(len(x), np.uint64)
    lib().or_compute_keys(len(x), _c(x, np.float64), _c(y, np.float64), _c(z, np.float64),
                          C.byref(b), keys)
    return keys


def hilbert_key(ix, iy, iz) -> int:
    return int(lib().or_intcoord_to_key_hilbert3d(int(ix), int(iy), int(iz)))


def key_to_intcoord(key) -> np.ndarray:
    ic = np.zeros(3, np.int64)
    lib().or_key_to_intcoord_hilbert(int(key), ic)
    return ic


def key_to_coord(b: Box, key) -> np.ndarray:
    x = np.zeros(3)
    lib().or_key_to_coord(C.byref(b), int(key), x)
    return x


def level_from_key(key) -> int:
    return int(lib().or_level_from_key(int(key)))


def sort_radix_ip(keys, idx=None):
    """Restated reference presort; returns (sorted keys, carried indices)."""
    k = np.array(keys, dtype=np.uint64)
    ix = np.arange(len(k), dtype=np.int64) if idx is None else np.array(idx, dtype=np.int64)
    lib().or_sort_radix_ip(len(k), k, ix, None)
    return k, ix


def dupfix(keys, x, y, z, b: Box):
    k = np.array(keys, dtype=np.uint64)
    xs, ys, zs = (np.array(v, dtype=np.float64) for v in (x, y, z))
    moved = lib().or_dupfix(len(k), k, xs, ys, zs, C.byref(b))
    if moved < 0:
        raise ValueError("key range cannot hold all particles")
    return k, xs, ys, zs, int(moved)


@dataclass
class Decomposition:
    box: Box
    keys: np.ndarray
    perm: np.ndarray
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    q: np.ndarray
    nfixed: int


def domain_decompose(x, y, z, q, force_cubic=False) -> Decomposition:
    n = len(x)
    b = Box()
    keys = np.empty(n, np.uint64)
    perm = np.empty(n, np.int64)
    xs, ys, zs, qs = (np.empty(n) for _ in range(4))
    nf = C.c_int64(0)
    rc = lib().or_domain_decompose(n, _c(x, np.float64), _c(y, np.float64), _c(z, np.float64),
                                   _c(q, np.float64), int(force_cubic), C.byref(b), keys, perm,
                                   xs, ys, zs, qs, C.byref(nf))
    if rc:
        raise ValueError(f"or_domain_decompose failed ({rc})")
    return Decomposition(b, keys, perm, xs, ys, zs, qs, int(nf.value))


@dataclass
class Tree:
    nnodes: int
    nleaf: int
    ntwig: int
    key: np.ndarray
    level: np.ndarray
    leaves: np.ndarray
    first: np.ndarray
    first_child: np.ndarray
    next_sibling: np.ndarray
    nchild: np.ndarray
    mom: np.ndarray
    node_leaf: np.ndarray
    b2: np.ndarray


def boxlength2(b: Box) -> np.ndarray:
    b2 = np.zeros(MAXLEVEL + 1)
    lib().or_boxlength2(C.byref(b), b2)
    return b2


def tree_build(keys, x, y, z, q, b: Box, cap=None) -> Tree:
    n = len(keys)
    cap = cap or (4 * n + 64)
    arr = dict(key=np.zeros(cap, np.uint64), level=np.zeros(cap, np.int32),
               leaves=np.zeros(cap, np.int64), first=np.zeros(cap, np.int64),
               first_child=np.zeros(cap, np.int64), next_sibling=np.zeros(cap, np.int64),
               nchild=np.zeros(cap, np.int32), mom=np.zeros(cap * MOM))
    node_leaf = np.full(n, -1, np.int64)
    counts = np.zeros(3, np.int64)
    rc = lib().or_tree_build(n, _c(keys, np.uint64), _c(x, np.float64), _c(y, np.float64),
                             _c(z, np.float64), _c(q, np.float64), cap,
                             arr["key"], arr["level"], arr["leaves"], arr["first"],
                             arr["first_child"], arr["next_sibling"], arr["nchild"], arr["mom"],
                             node_leaf, counts)
    if rc < 0:
        raise ValueError(f"or_tree_build failed ({rc})")
    nn = int(counts[0])
    out = {k: v[:nn] for k, v in arr.items() if k != "mom"}
    return Tree(nn, int(counts[1]), int(counts[2]), mom=arr["mom"][:nn * MOM].reshape(nn, MOM),
                node_leaf=node_leaf, b2=boxlength2(b), **out)


class mac_select:
    """Context manager: MAC of every oracle walk inside the block
    (coulomb/module_interaction_specific.f90:256-275): 0 Barnes-Hut, 1 Bmax
    min(bmax^2, 3 boxlength2), other values never accept (N^2)."""

    def __init__(self, value):
        self.value = int(value)

    def __enter__(self):
        self.prev = lib().or_get_mac_select()
        lib().or_set_mac_select(self.value)
        return self

    def __exit__(self, *exc):
        lib().or_set_mac_select(self.prev)
        return False


def walk_tiles(keys, tile_size, pack=0):
    """Walk tiles: maximal key cells with <= tile_size particles (reference
    rule); pack > 0 merges runs of sibling tiles up to pack particles."""
    n = len(keys)
    tf = np.zeros(n, np.int64)
    tc = np.zeros(n, np.int64)
    nt = lib().or_walk_tiles(n, _c(keys, np.uint64), int(tile_size), int(pack), tf, tc)
    if nt < 0:
        raise ValueError("insufficient key resolution in tile formation")
    return tf[:nt].copy(), tc[:nt].copy()


@dataclass
class WalkResult:
    e: np.ndarray          # (n, 3)
    pot: np.ndarray
    work: np.ndarray
    p2p: np.ndarray
    m2p: np.ndarray
    tile_first: np.ndarray
    tile_count: np.ndarray
    tile_mac: np.ndarray
    num_interactions: int
    num_mac: int
    ntiles_done: int


def walk_group(x, y, z, tree: Tree, theta2, eps2, tile_size=16, sign=1.0,
               tile_stride=1, nthreads=0, keys=None, tiles=None, pack=0) -> WalkResult:
    n = len(x)
    if tiles is None:
        tiles = walk_tiles(keys, tile_size, pack)
    tf, tc = tiles
    ex, ey, ez, pot, work = (np.zeros(n) for _ in range(5))
    p2p = np.zeros(n, np.int32)
    m2p = np.zeros(n, np.int32)
    tmac = np.zeros(len(tf), np.int64)
    tot = np.zeros(3, np.int64)
    rc = lib().or_walk_group(n, _c(x, np.float64), _c(y, np.float64), _c(z, np.float64),
                             tree.nnodes, tree.level, tree.leaves, tree.first_child,
                             tree.next_sibling, np.ascontiguousarray(tree.mom.ravel()), tree.b2,
                             float(theta2), float(eps2), float(sign), len(tf), tf, tc,
                             int(tile_stride), int(nthreads), ex, ey, ez, pot, work,
                             p2p, m2p, tmac, tot)
    if rc:
        raise ValueError("walk did not reach every leaf")
    return WalkResult(np.stack([ex, ey, ez], 1), pot, work, p2p, m2p, tf, tc, tmac,
                      int(tot[0]), int(tot[1]), int(tot[2]))


def p2p(charge, d, dist2):
    e = np.zeros(3)
    phi = np.zeros(1)
    lib().or_p2p(np.array([charge], np.float64), _c(d, np.float64), float(dist2), e, phi)
    return e, float(phi[0])


def m2p(mom, d, dist2):
    e = np.zeros(3)
    phi = np.zeros(1)
    lib().or_m2p(_c(mom, np.float64), _c(d, np.float64), float(dist2), e, phi)
    return e, float(phi[0])


def direct(x, y, z, q, testidx, eps2, sign=1.0, nthreads=0):
    t = _c(testidx, np.int64)
    m = len(t)
    ex, ey, ez, pot = (np.zeros(m) for _ in range(4))
    lib().or_direct(len(x), _c(x, np.float64), _c(y, np.float64), _c(z, np.float64),
                    _c(q, np.float64), m, t, float(eps2), float(sign), int(nthreads),
                    ex, ey, ez, pot)
    return np.stack([ex, ey, ez], 1), pot


# ------------------------------------------------- frozen lists and adjoint
@dataclass
class Lists:
    tile_first: np.ndarray
    tile_count: np.ndarray
    off: np.ndarray        # ntiles + 1
    node: np.ndarray       # entries in traversal order (leaf: P2P, twig: M2P)


def walk_lists(x, y, z, tree: Tree, theta2, tiles) -> Lists:
    """Interaction lists of the group walk (same decisions as walk_group)."""
    tf, tc = (_c(a, np.int64) for a in tiles)
    args = (_c(x, np.float64), _c(y, np.float64), _c(z, np.float64), tree.level,
            tree.first_child, tree.next_sibling, np.ascontiguousarray(tree.mom.ravel()), tree.b2,
            float(theta2), len(tf), tf, tc)
    need = lib().or_walk_lists(*args, 0, None, None)
    need = -need if need < 0 else need
    off = np.zeros(len(tf) + 1, np.int64)
    node = np.zeros(max(need, 1), np.int64)
    got = lib().or_walk_lists(*args, len(node), off.ctypes.data_as(C.c_void_p),
                              node.ctypes.data_as(C.c_void_p))
    assert got == need
    return Lists(tf, tc, off, node[:need])


def tree_moments(tree: Tree, x, y, z, q) -> np.ndarray:
    """Node moments of the same topology at other particle positions/charges."""
    mom = np.zeros(tree.nnodes * MOM)
    lib().or_tree_moments(tree.nnodes, tree.first, tree.first_child, tree.next_sibling,
                          _c(x, np.float64), _c(y, np.float64), _c(z, np.float64),
                          _c(q, np.float64), mom)
    return mom.reshape(tree.nnodes, MOM)


def eval_lists(x, y, z, tree: Tree, mom, lists: Lists, eps2, sign=1.0):
    n = len(x)
    ex, ey, ez, pot = (np.zeros(n) for _ in range(4))
    lib().or_eval_lists(_c(x, np.float64), _c(y, np.float64), _c(z, np.float64), tree.first,
                        tree.first_child, np.ascontiguousarray(np.asarray(mom).ravel()),
                        float(eps2), float(sign), len(lists.tile_first), lists.tile_first,
                        lists.tile_count, lists.off, lists.node, ex, ey, ez, pot)
    return np.stack([ex, ey, ez], 1), pot


def adjoint_term(d, r2, gphi, u, q, p=(0.0, 0.0, 0.0), qt=(0.0,) * 6):
    D, gq, gp, M = np.zeros(3), np.zeros(1), np.zeros(3), np.zeros(6)
    lib().or_adjoint_term(_c(d, np.float64), float(r2), float(gphi), _c(u, np.float64), float(q),
                          _c(p, np.float64), _c(qt, np.float64), D, gq, gp, M)
    return D, float(gq[0]), gp, M


def adjoint_lists(x, y, z, q, tree: Tree, lists: Lists, eps2, ge, gpot, sign=1.0):
    """dL/d(x, y, z, q) of L = sum gpot.pot + ge.E over frozen lists."""
    n = len(x)
    ge = np.asarray(ge, np.float64)
    gx, gy, gz, gq = (np.zeros(n) for _ in range(4))
    lib().or_adjoint_lists(n, _c(x, np.float64), _c(y, np.float64), _c(z, np.float64),
                           _c(q, np.float64), tree.nnodes, tree.first, tree.leaves,
                           tree.first_child, np.ascontiguousarray(tree.mom.ravel()), float(eps2),
                           float(sign), len(lists.tile_first), lists.tile_first, lists.tile_count,
                           lists.off, lists.node, _c(ge[:, 0], np.float64),
                           _c(ge[:, 1], np.float64), _c(ge[:, 2], np.float64),
                           _c(gpot, np.float64), gx, gy, gz, gq)
    return gx, gy, gz, gq


def solve(x, y, z, q, theta2, eps2, tile_size=16, sign=1.0, nthreads=0, force_cubic=False,
          pack=0):
    """Full single-rank grow + traverse as the reference frontends call it.

    Returns (decomposition, tree, walk result); results are in sorted order.
    """
    d = domain_decompose(x, y, z, q, force_cubic)
    t = tree_build(d.keys, d.x, d.y, d.z, d.q, d.box)
    w = walk_group(d.x, d.y, d.z, t, theta2, eps2, tile_size=tile_size, sign=sign,
                   nthreads=nthreads, keys=d.keys, pack=pack)
    return d, t, w


# ---------------------------------------------------------------- reference sort
_ref = None


def ref_lib():
    """The reference's own sort_radix_ip (oracle/_ref), or None if not built."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB_PATH):
        L = C.CDLL(REF_LIB_PATH)
        L.ref_sort_radix_ip.argtypes = [C.c_longlong, _i64, _i64, _f64]
        L.ref_sort_radix_ip.restype = C.c_int
        L.ref_sizeof_key.restype = C.c_int
        L.ref_sizeof_index.restype = C.c_int
        _ref = L
    return _ref


def ref_sort_radix_ip(keys):
    L = ref_lib()
    if L is None:
        raise RuntimeError("oracle/_ref/libref_sort.so not built")
    k = np.array(keys, dtype=np.uint64).view(np.int64).copy()
    ix = np.arange(1, len(k) + 1, dtype=np.int64)   # indxl init, mpi_pepckeys.c:161
    w = np.ones(len(k))
    rc = L.ref_sort_radix_ip(len(k), k, ix, w)
    if rc:
        raise RuntimeError(f"reference sort failed ({rc})")
    return k.view(np.uint64), ix - 1
