"""Seeded synthetic particle systems for the Coulomb tree solve.

Each generator returns float64 numpy arrays (positions, charges, masses,
velocities) plus the solver parameters of that workload.  The shapes and
physics follow the five benchmark systems of this project:

* ``mini``       32,768 charges in the unit cube, alternating +1/-1 charges
                 (as pepc-mini labels them, src/frontends/pepc-mini/
                 module_helper.f90:185), eps=0.01, theta=0.6
* ``explosion``  4,194,304 equal positive ions q=1/N uniform in the unit
                 ball at rest (the Coulomb explosion of pepc-benchmark,
                 src/frontends/pepc-benchmark/module_helper.f90:195-233),
                 eps=1e-3, theta in {0.3, 0.6, 0.9}
* ``plummer``    8,388,608 equal masses: Plummer sphere (a=1, r<=20) plus a
                 10% uniform background in [-20,20]^3, attractive sign,
                 eps=1e-3, theta=0.5
* ``ensemble``   256 independent neutral cubes, N_s log-uniform in
                 [1024, 65536], Maxwellian velocities, eps=0.01, theta=0.6
* ``adjoint``    262,144 charges q~N(0,1) in the unit cube, eps=0.01,
                 theta=0.6, targets from 1%-perturbed charges

All use numpy's PCG64 generator with the recorded seed; data is synthetic.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class System:
    name: str
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    q: np.ndarray
    m: np.ndarray
    v: np.ndarray           # (n, 3)
    eps2: float
    theta2: float
    sign: float = 1.0       # +1 Coulomb (repulsive like charges), -1 gravity
    dt: float = 0.0
    nsteps: int = 0
    seed: int = 0
    extra: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return len(self.x)


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def mini(n=32768, seed=1):
    r = _rng(seed)
    p = r.random((n, 3))
    label = np.arange(1, n + 1)
    q = -1.0 + 2.0 * np.mod(label, 2)
    return System("mini", p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), q,
                  np.ones(n), np.zeros((n, 3)), eps2=1e-4, theta2=0.36, dt=0.01,
                  nsteps=10, seed=seed)


def periodic(n=32768, seed=11, box=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0), eps2=1e-4,
             theta2=0.36):
    """Neutral plasma in a periodic box: uniform positions, alternating
    +-1 charges (pepc-mini's charges), unit masses, at rest."""
    r = _rng(seed)
    p = np.asarray(origin, float) + r.random((n, 3)) * np.asarray(box, float)
    q = -1.0 + 2.0 * np.mod(np.arange(1, n + 1), 2)
    return System("periodic", p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), q,
                  np.ones(n), np.zeros((n, 3)), eps2=eps2, theta2=theta2, dt=0.002,
                  nsteps=10, seed=seed, extra={"box": tuple(box), "origin": tuple(origin)})


def unit_ball(r, n):
    u = r.random(n)
    rad = np.cbrt(u)
    v = r.standard_normal((n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    return v * rad[:, None]


def explosion(n=4194304, theta=0.6, seed=2):
    r = _rng(seed)
    p = unit_ball(r, n)
    return System("explosion", p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(),
                  np.full(n, 1.0 / n), np.ones(n), np.zeros((n, 3)), eps2=1e-6,
                  theta2=theta * theta, dt=1e-3, nsteps=20, seed=seed)


def plummer(n=8388608, seed=3, theta=0.5):
    r = _rng(seed)
    nb = int(round(0.1 * n))
    npl = n - nb
    rad = np.empty(npl)
    filled = 0
    while filled < npl:   # inverse CDF of M(r)=r^3/(1+r^2)^1.5, reject r > 20
        u = r.random(npl - filled)
        u = np.clip(u, 1e-300, 1.0)
        rr = 1.0 / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
        rr = rr[rr <= 20.0]
        rad[filled:filled + len(rr)] = rr
        filled += len(rr)
    d = r.standard_normal((npl, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    p = np.concatenate([d * rad[:, None], r.uniform(-20.0, 20.0, (nb, 3))])
    return System("plummer", p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(),
                  np.full(n, 1.0 / n), np.full(n, 1.0 / n), np.zeros((n, 3)), eps2=1e-6,
                  theta2=theta * theta, sign=-1.0, seed=seed)


def ensemble(nsys=256, seed=4, nmin=1024, nmax=65536):
    r = _rng(seed)
    sizes = np.floor(np.exp(r.uniform(np.log(nmin), np.log(nmax), nsys))).astype(np.int64)
    systems = []
    for s in range(nsys):
        ns = int(sizes[s])
        side = ns ** (1.0 / 3.0) / 10.0
        p = r.random((ns, 3)) * side
        q = -1.0 + 2.0 * np.mod(np.arange(1, ns + 1), 2)
        v = r.standard_normal((ns, 3))
        systems.append(System(f"ensemble[{s}]", p[:, 0].copy(), p[:, 1].copy(),
                              p[:, 2].copy(), q, np.ones(ns), v, eps2=1e-4, theta2=0.36,
                              dt=1e-3, nsteps=50, seed=seed))
    return systems


def concat(systems):
    """Concatenates independent systems into one batched input: returns
    (x, y, z, q, m, v, offsets) with system s in [offsets[s], offsets[s+1])."""
    off = np.zeros(len(systems) + 1, np.int64)
    off[1:] = np.cumsum([s.n for s in systems])
    cat = lambda name: np.concatenate([getattr(s, name) for s in systems])
    return cat("x"), cat("y"), cat("z"), cat("q"), cat("m"), cat("v"), off


def adjoint(n=262144, seed=5):
    r = _rng(seed)
    p = r.random((n, 3))
    q = r.standard_normal(n)
    qt = q + 0.01 * r.standard_normal(n) * q
    return System("adjoint", p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), q,
                  np.ones(n), np.zeros((n, 3)), eps2=1e-4, theta2=0.36, seed=seed,
                  extra={"q_target": qt})


def sample_targets(n, m, seed=7):
    """m distinct target indices out of n (without replacement), sorted."""
    r = _rng(seed)
    m = min(m, n)
    return np.sort(r.choice(n, size=m, replace=False)).astype(np.int64)


def coincident(n=20000, seed=8, pairs=0.05, triples=0.01, cluster=0, cluster_side=2e-4,
               cluster_at=(0.3, 0.3, 0.3)):
    """Particles that share positions, so that keys tie: uniform in the unit
    cube, optionally `cluster` of them in a small cube of side cluster_side
    (its particles share many leading key digits), then a fraction `pairs`
    of the particles copied onto other particles' positions and a fraction
    `triples` placed on a third copy; random +-1 charges, eps=0.01,
    theta=0.6.  Exercises the order of equal keys in the sort and the
    duplicate-key fix-up that separates them."""
    r = _rng(seed)
    p = r.random((n, 3))
    if cluster:
        p[:cluster] = np.asarray(cluster_at) + cluster_side * r.random((cluster, 3))
    idx = r.permutation(n)
    npair = int(pairs * n)
    ntri = int(triples * n)
    src, dst = idx[:npair], idx[npair:2 * npair]
    p[dst] = p[src]
    third = idx[2 * npair:2 * npair + ntri]
    p[third] = p[src[:ntri]]
    q = r.choice(np.array([-1.0, 1.0]), n)
    return System("coincident", p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), q,
                  np.ones(n), np.zeros((n, 3)), eps2=1e-4, theta2=0.36, seed=seed,
                  extra={"pairs": npair, "triples": ntri, "cluster": cluster})
