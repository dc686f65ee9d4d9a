"""Periodic boundaries on the GPU (through the C-ABI) against the oracle.

Near field: the group walk over the mirror boxes must reproduce the
oracle's interaction counts per target exactly (MAC decisions on shifted
targets) and its fields to 1e-12 in FP64.  Far field: the lattice
coefficients are the reference's table bit for bit for the unit cube and
match the oracle's computation elsewhere; mu (the lattice's local expansion)
and the total fields match the oracle to 1e-11 (same formulas in a different
evaluation order: Cartesian recurrences instead of Legendre calls).  The
physics checks (Ewald sum, Madelung constant) hold for the GPU output
itself.
"""
import json
import math
import os

import numpy as np
import pytest

from pepc_amd import workloads

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pepc_amd.solver import Solver
    return Solver


@pytest.fixture(scope="module")
def P():
    from oracle import periodic
    return periodic


@pytest.fixture(scope="module")
def table(P):
    with open(os.path.join(HERE, "golden", "lattice_unit_cube.json")) as f:
        return P.unit_cube_table(json.load(f))


def _rows_close(a, b, rows, rtol, lmax):
    """Every entry of rows l in `rows`: |a - b| <= rtol * S_l, S_l the largest
    |b| of rows l-2..l+2 (row 0 only for itself).  Rows that a symmetry of the
    lattice makes vanish hold rounding noise; their neighbours set the scale."""
    from oracle.periodic import idx
    row = lambda l: np.abs(b[idx(l, 0):idx(l, l) + 1]).max()
    for l in rows:
        near = [l] if l == 0 else [l, *range(max(2, l - 2), min(lmax, l + 2) + 1)]
        scale = max(row(k) for k in near)
        sl = slice(idx(l, 0), idx(l, l) + 1)
        err = np.abs(a[sl] - b[sl]).max()
        assert err <= rtol * scale, (l, err, scale)


def test_unit_cube_table_bitexact(gpu, table):
    sol = gpu(precision="fp64")
    sol.set_periodic()
    ml, _, _ = sol.lattice()
    assert np.array_equal(ml, table)
    sol.close()


@pytest.mark.parametrize("lattice", [(1.3, 1.3, 1.3), (1.0, 1.5, 0.8)])
def test_computed_coefficients(gpu, P, lattice):
    sol = gpu(precision="fp64")
    sol.set_periodic(lattice=lattice)
    ml, _, _ = sol.lattice()
    ref = P.lattice_coefficients(lattice=lattice)
    print(f"[coefficients {lattice}] ML00 {ml[0].real:.6e} / {ref[0].real:.6e}", flush=True)
    # row 1 is rounding noise amplified 3^32 times by the reference's iteration
    _rows_close(ml, ref, [0] + list(range(2, 41)), 1e-11, 40)
    sol.close()


def _solve(gpu, s, precision, G=16, **per):
    sol = gpu(theta2=s.theta2, eps2=s.eps2, group_size=G, precision=precision)
    sol.set_periodic(**per)
    st = sol.grow(s.x, s.y, s.z, s.q)
    res = sol.traverse()
    return sol, st, res


def _oracle(P, s, ml, box, origin, layers=1, far=True, periodicity=(1, 1, 1), G=16,
            correction=2):
    from oracle import oracle as orc
    d = orc.domain_decompose(s.x, s.y, s.z, s.q)
    t = orc.tree_build(d.keys, d.x, d.y, d.z, d.q, d.box)
    vb = P.lattice_vectors(box, P.neighbour_boxes(periodicity, layers))
    w = P.walk_group_boxes(d.x, d.y, d.z, t, s.theta2, s.eps2, vb, tile_size=G, keys=d.keys)
    lat = None
    if far:
        lat = P.Lattice(ml, box, origin, layers=layers, correction=correction)
        lat.timestep(d.x, d.y, d.z, d.q)
        e, p = lat.field(d.x, d.y, d.z)
        w.e = w.e + e
        w.pot = w.pot + p
    return d, w, lat


CASES = {
    "unit": dict(n=4096, box=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0)),
    "box2": dict(n=4096, box=(2.0, 2.0, 2.0), origin=(-1.0, -1.0, -1.0)),
    "rect": dict(n=3000, box=(1.0, 1.5, 0.8), origin=(0.25, -0.5, 0.0)),
}


@pytest.mark.parametrize("name", list(CASES))
def test_periodic_solve_fp64_parity(gpu, P, table, name):
    c = CASES[name]
    s = workloads.periodic(n=c["n"], seed=31, box=c["box"], origin=c["origin"])
    sol, st, res = _solve(gpu, s, "fp64", lattice=c["box"], origin=c["origin"])
    ml = sol.lattice()[0] if name != "unit" else table
    d, w, lat = _oracle(P, s, ml, c["box"], c["origin"])
    keys, perm, *_ = sol.sorted_host()
    assert np.array_equal(perm, d.perm)
    wk = sol.read_walk()
    assert np.array_equal(wk["p2p"].astype(np.int32), w.p2p)
    assert np.array_equal(wk["m2p"].astype(np.int32), w.m2p)
    assert sol.sync()["interactions"] == int(w.p2p.sum() + w.m2p.sum())
    _, mu, fict = sol.lattice()
    _rows_close(mu, lat.step.mu[:231], range(21), 1e-11, 20)
    assert np.abs(fict - lat.step.fict).max() <= 1e-13 * max(1.0, np.abs(lat.step.fict).max())
    e = res.e.cpu().numpy().T
    pot = res.pot.cpu().numpy()
    scale = np.abs(w.e).max()
    print(f"[periodic {name}] n={s.n} p2p={w.p2p.sum()} m2p={w.m2p.sum()} "
          f"max|dE|/max|E|={np.abs(e - w.e).max() / scale:.2e}", flush=True)
    assert np.abs(e - w.e).max() <= 1e-11 * scale
    assert np.abs(pot - w.pot).max() <= 1e-11 * np.abs(w.pot).max()
    sol.close()


@pytest.mark.parametrize("name", ["unit", "box2"])
@pytest.mark.parametrize("correction", [0, 1])
def test_extrinsic_corrections_fp64_parity(gpu, P, table, name, correction):
    """fmm_extrinsic_correction NONE (0) and REDLACK (1) (module_fmm_framework.
    f90:807-829) against the oracle under the same setting: counts exact, the
    field to 1e-11, the potential to 1e-11 in the unit cube and up to a
    constant in a computed lattice (the constant carries the l = 1 lattice
    row's rounding noise times the box dipole, which the two computations do
    not share)."""
    c = CASES[name]
    s = workloads.periodic(n=c["n"], seed=38, box=c["box"], origin=c["origin"])
    sol, st, res = _solve(gpu, s, "fp64", lattice=c["box"], origin=c["origin"],
                          correction=correction)
    ml = sol.lattice()[0] if name != "unit" else table
    d, w, lat = _oracle(P, s, ml, c["box"], c["origin"], correction=correction)
    wk = sol.read_walk()
    assert np.array_equal(wk["p2p"].astype(np.int32), w.p2p)
    assert np.array_equal(wk["m2p"].astype(np.int32), w.m2p)
    _, _, fict = sol.lattice()
    assert len(fict) == 0 and len(lat.step.fict) == 0
    e = res.e.cpu().numpy().T
    pot = res.pot.cpu().numpy()
    scale = np.abs(w.e).max()
    print(f"[correction {correction} {name}] max|dE|/max|E|={np.abs(e - w.e).max() / scale:.2e}",
          flush=True)
    assert np.abs(e - w.e).max() <= 1e-11 * scale
    dp = pot - w.pot
    if name == "unit":
        assert np.abs(dp).max() <= 1e-11 * np.abs(w.pot).max()
    else:
        assert np.abs(dp - dp.mean()).max() <= 1e-11 * np.abs(w.pot).max()
    sol.close()


def test_redlack_equals_ewald_on_gpu(gpu, P):
    """theta = 0, unit cube, Redlack-Grindlay: the GPU sum is the tin-foil
    Ewald sum itself, field and potential (no constant), to the scheme's
    truncation at multipole order 20."""
    s = workloads.periodic(n=64, seed=34, eps2=0.0, theta2=0.0)
    sol, st, res = _solve(gpu, s, "fp64", correction=1)
    keys, perm, xs, ys, zs, qs = sol.sorted_host()
    e = res.e.cpu().numpy().T
    pot = res.pot.cpu().numpy()
    e_ew, p_ew = P.ewald(xs, ys, zs, qs)
    rms = math.sqrt((e_ew ** 2).mean())
    prms = math.sqrt((p_ew ** 2).mean())
    print(f"[ewald redlack] max|dE|/rms {np.abs(e - e_ew).max() / rms:.2e} "
          f"max|dphi|/rms {np.abs(pot - p_ew).max() / prms:.2e}", flush=True)
    assert np.abs(e - e_ew).max() <= 1e-5 * rms
    assert np.abs(pot - p_ew).max() <= 1e-5 * prms
    sol.close()


def test_two_mirror_layers_far_field(gpu, P):
    """mirror_box_layers = 2 with the far field (ws = 2): the coefficients
    computed on the GPU (double-double sums of the 124 x 125 L* terms per
    entry) against the oracle's sorted sums to 1e-11 of the row scale (row 1
    excluded, rounding noise); the solve over 125 mirror boxes against the
    oracle (counts exact, field 1e-11, potential 1e-11 up to a constant) with
    the GPU's coefficients; and, at
    theta = 0, the GPU output against the tin-foil Ewald sum (field 1e-7 of
    its rms, potential up to -(2 pi / 3V) sum q r^2)."""
    ref = P.lattice_coefficients(layers=2)
    s = workloads.periodic(n=2000, seed=39)
    sol, st, res = _solve(gpu, s, "fp64", layers=2)
    ml, _, _ = sol.lattice()
    _rows_close(ml, ref, [0] + list(range(2, 41)), 1e-11, 40)
    d, w, lat = _oracle(P, s, ml, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), layers=2)
    wk = sol.read_walk()
    assert np.array_equal(wk["p2p"].astype(np.int32), w.p2p)
    assert np.array_equal(wk["m2p"].astype(np.int32), w.m2p)
    e = res.e.cpu().numpy().T
    pot = res.pot.cpu().numpy()
    scale = np.abs(w.e).max()
    dp = pot - w.pot
    print(f"[two layers] p2p={w.p2p.sum()} m2p={w.m2p.sum()} "
          f"max|dE|/max|E|={np.abs(e - w.e).max() / scale:.2e} dphi const {dp.mean():.2e}", flush=True)
    assert np.abs(e - w.e).max() <= 1e-11 * scale
    # up to a constant: the l = 1 lattice row's rounding noise times the
    # (rounding-level) remaining dipole of the box
    assert np.abs(dp - dp.mean()).max() <= 1e-11 * np.abs(w.pot).max()
    sol.close()
    s = workloads.periodic(n=64, seed=34, eps2=0.0, theta2=0.0)
    sol, st, res = _solve(gpu, s, "fp64", layers=2)
    keys, perm, xs, ys, zs, qs = sol.sorted_host()
    e = res.e.cpu().numpy().T
    pot = res.pot.cpu().numpy()
    e_ew, p_ew = P.ewald(xs, ys, zs, qs)
    rms = math.sqrt((e_ew ** 2).mean())
    print(f"[two layers ewald] max|dE|/rms {np.abs(e - e_ew).max() / rms:.2e}", flush=True)
    assert np.abs(e - e_ew).max() <= 1e-7 * rms
    r2 = (xs - 0.5) ** 2 + (ys - 0.5) ** 2 + (zs - 0.5) ** 2
    shift = -2 * math.pi / 3 * (qs * r2).sum()
    assert np.abs(pot - p_ew - shift).max() <= 1e-7 * math.sqrt((p_ew ** 2).mean())
    sol.close()


def test_periodic_solve_fp32(gpu, P, table):
    c = CASES["unit"]
    s = workloads.periodic(n=c["n"], seed=32)
    sol, st, res = _solve(gpu, s, "fp32")
    d, w, _ = _oracle(P, s, table, c["box"], c["origin"])
    wk = sol.read_walk()
    assert np.array_equal(wk["p2p"].astype(np.int32), w.p2p)
    assert np.array_equal(wk["m2p"].astype(np.int32), w.m2p)
    e = res.e.cpu().numpy().T
    rel = np.linalg.norm(e - w.e, axis=1) / np.linalg.norm(w.e, axis=1)
    print(f"[periodic fp32] median {np.median(rel):.2e} max {rel.max():.2e}", flush=True)
    assert np.median(rel) < 1e-6 and rel.max() < 1e-4
    sol.close()


def test_near_field_partial_periodicity(gpu, P):
    s = workloads.periodic(n=3000, seed=33)
    sol, st, res = _solve(gpu, s, "fp64", periodicity=(1, 1, 0), layers=2, far_field=False)
    d, w, _ = _oracle(P, s, None, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), layers=2, far=False,
                      periodicity=(1, 1, 0))
    wk = sol.read_walk()
    assert np.array_equal(wk["p2p"].astype(np.int32), w.p2p)
    assert np.array_equal(wk["m2p"].astype(np.int32), w.m2p)
    e = res.e.cpu().numpy().T
    assert np.abs(e - w.e).max() <= 1e-12 * np.abs(w.e).max()
    sol.close()


def test_ewald_and_madelung_on_gpu(gpu, P):
    """theta = 0: the GPU sum itself against the Ewald sum and the rock-salt
    Madelung constant (no oracle involved).  Tolerances are the scheme's own
    truncation at multipole order 20 (largest for particles near the box
    corners), which the oracle shares."""
    s = workloads.periodic(n=64, seed=34, eps2=0.0, theta2=0.0)
    sol, st, res = _solve(gpu, s, "fp64")
    keys, perm, xs, ys, zs, qs = sol.sorted_host()
    e = res.e.cpu().numpy().T
    pot = res.pot.cpu().numpy()
    e_ew, p_ew = P.ewald(xs, ys, zs, qs)
    rms = math.sqrt((e_ew ** 2).mean())
    print(f"[ewald] max|dE|/rms {np.abs(e - e_ew).max() / rms:.2e}", flush=True)
    assert np.abs(e - e_ew).max() <= 1e-5 * rms
    r2 = (xs - 0.5) ** 2 + (ys - 0.5) ** 2 + (zs - 0.5) ** 2
    shift = -2 * math.pi / 3 * (qs * r2).sum()
    assert np.abs(pot - p_ew - shift).max() <= 1e-5 * math.sqrt((p_ew ** 2).mean())
    u, u_ew = 0.5 * (qs * pot).sum(), 0.5 * (qs * p_ew).sum()
    assert abs(u - u_ew) <= 1e-6 * abs(u_ew)
    sol.close()
    x, y, z, q = P.rock_salt(2)
    x, y, z = x + 0.0625, y + 0.0625, z + 0.0625
    sol = gpu(theta2=0.0, eps2=0.0, precision="fp64")
    sol.set_periodic()
    sol.grow(x, y, z, q)
    res = sol.traverse()
    _, _, _, _, _, qs = sol.sorted_host()
    madelung = -res.pot.cpu().numpy() * qs * 0.25     # nearest-neighbour distance 1/4
    print(f"[madelung] {madelung.min():.9f} .. {madelung.max():.9f}", flush=True)
    assert np.abs(madelung - P.MADELUNG_NACL).max() < 2e-7
    sol.close()


def test_constrain_periodic(gpu):
    import torch
    sol = gpu(precision="fp64")
    sol.set_periodic(lattice=(1.0, 2.0, 0.5), origin=(-0.5, 0.0, 1.0), periodicity=(1, 1, 0),
                     far_field=False)
    r = np.random.default_rng(3)
    x, y, z = (r.uniform(-3, 3, 1000) for _ in range(3))
    tx, ty, tz = (torch.from_numpy(v.copy()).cuda() for v in (x, y, z))
    sol.constrain_periodic(tx, ty, tz)
    wx = np.mod(x + 0.5, 1.0) - 0.5
    wy = np.mod(y - 0.0, 2.0) + 0.0
    assert np.array_equal(tx.cpu().numpy(), ((x + 0.5) - np.floor((x + 0.5) / 1.0) * 1.0) - 0.5)
    assert np.array_equal(ty.cpu().numpy(), ((y - 0.0) - np.floor((y - 0.0) / 2.0) * 2.0) + 0.0)
    assert np.array_equal(tz.cpu().numpy(), z)
    assert np.allclose(tx.cpu().numpy(), wx) and np.allclose(ty.cpu().numpy(), wy)
    sol.close()


def test_periodic_timestep_energy_parity(gpu, P, table):
    """Five on-device leapfrog steps with periodic boundaries: positions stay
    in the box (wrapped after each drift), and each step's energy record is
    1/2 sum q phi of the oracle's periodic solve of that step's state."""
    s = workloads.periodic(n=4096, seed=35)
    sol = gpu(theta2=s.theta2, eps2=s.eps2, precision="fp64")
    sol.set_periodic()
    sol.state_set(s.x, s.y, s.z, s.v, s.q, s.m)
    for step in range(6):
        st = sol.state_get()
        pos = np.stack([st["x"], st["y"], st["z"]], 1)
        assert (pos >= 0.0).all() and (pos <= 1.0).all()
        ss = workloads.System("p", st["x"], st["y"], st["z"], st["q"], st["m"], st["v"],
                              eps2=s.eps2, theta2=s.theta2)
        d, w, _ = _oracle(P, ss, table, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0))
        en = sol.energies()
        epot = 0.5 * np.sum(d.q * w.pot)
        print(f"[periodic step {step}] epot {en[step, 0, 0]:.12e} oracle {epot:.12e}", flush=True)
        assert abs(en[step, 0, 0] - epot) <= 1e-10 * abs(epot)
        if step < 5:
            sol.run(1, s.dt)
    sol.close()


def test_periodic_graph_replay_equals_eager(gpu):
    """The periodic step (wrap, mirror walks, lattice kernels) captured in a
    HIP graph and replayed gives the eager steps' energies bit for bit."""
    s = workloads.periodic(n=3000, seed=36)
    runs = []
    for graph in (False, True):
        sol = gpu(theta2=s.theta2, eps2=s.eps2, precision="fp32", group_size=32)
        sol.set_periodic()
        sol.state_set(s.x, s.y, s.z, s.v, s.q, s.m)
        for _ in range(3):
            sol.run(1, s.dt, graph=graph)
        runs.append((sol.energies(), sol.state_get()))
        sol.close()
    (e0, st0), (e1, st1) = runs
    assert np.array_equal(e0, e1)
    assert np.array_equal(st0["x"], st1["x"]) and np.array_equal(st0["v"], st1["v"])


def test_periodic_errors(gpu):
    from pepc_amd._lib import PepcGpuError
    sol = gpu(precision="fp64")
    sol.set_periodic(periodicity=(1, 1, 0))       # the far field runs with two periodic axes too
    assert np.isfinite(sol.lattice()[0][1:]).all()
    with pytest.raises(PepcGpuError):
        sol.set_periodic(correction=3)                           # MEASUREMENT: not in the reference
    with pytest.raises(PepcGpuError):
        sol.set_periodic(layers=4, far_field=False)
    with pytest.raises(PepcGpuError):
        sol.set_periodic(lattice=np.zeros((3, 3)))               # degenerate lattice
    s = workloads.periodic(n=2000, seed=37)
    sol.set_periodic()
    sol.grow(s.x, s.y, s.z, s.q)
    sol.traverse()
    with pytest.raises(PepcGpuError):
        sol.freeze()                                             # no frozen lists with periodic
    import torch
    g = torch.zeros((3, s.n), dtype=torch.float64, device="cuda")
    with pytest.raises(PepcGpuError):
        sol.adjoint(g, g[0])
    sol.set_periodic(periodicity=None)                           # cleared: plain solve again
    sol.grow(s.x, s.y, s.z, s.q)
    sol.traverse()
    sol.close()
