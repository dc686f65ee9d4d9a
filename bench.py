#!/usr/bin/env python3
"""Benchmark: Coulomb particle updates/s of the Barnes-Hut solve on one GPU.

Workload: the Coulomb explosion system, N = 4,194,304 equal positive ions
(q = 1/N) uniform in the unit ball, eps = 1e-3, theta = 0.6, quadrupole
moments; FP32 + Kahan kernels (FP64 upward pass and acceptance tests).
A step is one grow + traverse of the whole system (keys, sort, tree build,
upward pass, walk and P2P/M2P) with the particles already resident in HBM.

Reported beside the throughput: the relative force error against an FP64
direct sum on sampled targets, the same solve on the FP64 kernels, the
roofline position of the walk kernel (the dominant kernel) and the CPU
restatement of the reference algorithm timed on this host.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_VALU_TFLOPS = 157.3     # MI355X published FP32 vector peak
# Wave64 VALU instruction issue implied by that peak: 157.3e12 / (64 lanes x 2 flops per FMA)
PEAK_VALU_ISSUE = PEAK_FP32_VALU_TFLOPS * 1e12 / 128.0
ERROR_BAR = 1e-4                  # the metric's relative force error bar
FLOPS_P2P, FLOPS_M2P, FLOPS_MAC = 22, 135, 9   # per interaction, reference formulas


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def walk_flops(st):
    return FLOPS_P2P * st["p2p"] + FLOPS_M2P * st["m2p"] + FLOPS_MAC * st["mac_evaluations"]


def run_gpu(sys_, precision, group, steps, warmup, dev_inputs, pack=0):
    import torch

    from pepc_amd.solver import Solver
    sol = Solver(theta2=sys_.theta2, eps2=sys_.eps2, group_size=group, precision=precision,
                 sign=sys_.sign, group_pack=pack)
    x, y, z, q = dev_inputs
    out = None
    for _ in range(warmup):
        sol.grow(x, y, z, q, sync=False)
        out = sol.traverse(out, sync=False)
    st = sol.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sol.grow(x, y, z, q, sync=False)
        out = sol.traverse(out, sync=False)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    st = sol.sync()
    ms = (t1 - t0) * 1e3 / steps
    # per-kernel device times (HIP events on the context stream), outside the timed region
    tg, tw = [], []
    for _ in range(min(steps, 5)):
        sol.grow(x, y, z, q, sync=False)
        out = sol.traverse(out, sync=False)
        g, w = sol.timings()
        tg.append(g)
        tw.append(w)
    return sol, out, st, ms, float(np.mean(tg)), float(np.mean(tw))


def run_timestep(sys_, group, pack, steps, dt=1e-3):
    """The same solve inside the on-device leapfrog step (graph replay): each
    step pushes the particles and grows from the previous step's key order,
    as a frontend's time loop does."""
    import torch

    from pepc_amd.solver import Solver
    sol = Solver(theta2=sys_.theta2, eps2=sys_.eps2, group_size=group, sign=sys_.sign,
                 group_pack=pack)
    sol.state_set(sys_.x, sys_.y, sys_.z, sys_.v, sys_.q, sys_.m, max_records=steps + 4)
    sol.run(2, dt, graph=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sol.run(steps, dt, graph=True, sync=False)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    sol.sync()
    en = sol.energies()
    sol.close()
    tot = en[:, 0, 0] + en[:, 0, 1]
    return ms, float(abs(tot[-1] - tot[0]) / abs(tot[0]))


def _timed(fn, reps):
    """Best device-synchronised wall time of fn() over reps runs (ms)."""
    import torch
    best = None
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) * 1e3
        best = t if best is None else min(best, t)
    return best


def run_other_configs(steps, warmup):
    """The other configurations of the solve, one record each (parity-test
    cases, timed here for the measured-results table): the small FP64 system
    with its 10 leapfrog steps, the clustered 8M gravity system, the batched
    ensemble of 256 systems with its 50-step graph timestep, and the adjoint
    (forward + backward through torch.autograd)."""
    import torch

    from pepc_amd import workloads
    from pepc_amd.autograd import coulomb_field
    from pepc_amd.solver import Solver
    dev = torch.device("cuda", 0)
    recs = []

    # small system, FP64: one solve, then 10 leapfrog steps dt = 0.01
    s = workloads.mini()
    X = [torch.from_numpy(a).to(dev) for a in (s.x, s.y, s.z, s.q)]
    sol = Solver(theta2=s.theta2, eps2=s.eps2, group_size=16, precision="fp64", sign=s.sign)
    for _ in range(warmup):
        sol.grow(*X, sync=False)
        sol.traverse(sync=False)
    ms = _timed(lambda: (sol.grow(*X, sync=False), sol.traverse(sync=False)), steps)
    sol.state_set(s.x, s.y, s.z, s.v, s.q, s.m, max_records=16)
    sol.run(1, 0.01, graph=True)
    ms_step = _timed(lambda: sol.run(10, 0.01, graph=True, sync=False), 3) / 10
    sol.close()
    recs.append({"workload": "small_n32768_fp64", "n": s.n, "precision": "fp64", "group_size": 16,
                 "solve_ms": ms, "updates_per_s": s.n / (ms * 1e-3),
                 "leapfrog_ms_per_step": ms_step})

    # clustered gravity, 8,388,608 particles, attractive sign, theta = 0.5
    s = workloads.plummer()
    X = [torch.from_numpy(a).to(dev) for a in (s.x, s.y, s.z, s.q)]
    for g, pk in ((16, 32), (64, 64)):
        sol = Solver(theta2=s.theta2, eps2=s.eps2, group_size=g, group_pack=pk, precision="fp32",
                     sign=s.sign)
        for _ in range(warmup):
            sol.grow(*X, sync=False)
            sol.traverse(sync=False)
        ms = _timed(lambda: (sol.grow(*X, sync=False), sol.traverse(sync=False)), steps)
        st = sol.sync()
        recs.append({"workload": "plummer_n8388608_theta0.5", "n": s.n, "precision": "fp32",
                     "group_size": g, "group_pack": pk, "solve_ms": ms,
                     "updates_per_s": s.n / (ms * 1e-3), "interactions": int(st["interactions"]),
                     "maxlevel": int(st["maxlevel"])})
        sol.close()
    del X

    # batched ensemble: 256 systems, autotuned rule, one solve and the 50-step graph timestep
    systems = workloads.ensemble()
    x, y, z, q, m, v, off = workloads.concat(systems)
    n = int(off[-1])
    sol = Solver(theta2=0.36, eps2=1e-4, precision="fp32")
    best, _ = sol.autotune(x, y, z, q, offsets=off, reps=1)
    X = [torch.from_numpy(a).to(dev) for a in (x, y, z, q)]
    for _ in range(warmup):
        sol.grow_batch(*X, off, sync=False)
        sol.traverse(sync=False)
    ms = _timed(lambda: (sol.grow_batch(*X, off, sync=False), sol.traverse(sync=False)), steps)
    sol.state_set(x, y, z, v, q, m, offsets=off, max_records=64)
    sol.run(1, 1e-3, graph=True)
    t50 = _timed(lambda: sol.run(50, 1e-3, graph=True, sync=False), 1)
    cfg = sol.config()
    sol.close()
    recs.append({"workload": "ensemble_256_systems", "n": n, "systems": len(systems),
                 "precision": "fp32", "group_size": cfg["group_size"],
                 "group_pack": cfg["group_pack"], "solve_ms": ms,
                 "updates_per_s": n / (ms * 1e-3), "timestep_50_ms": t50,
                 "timestep_updates_per_s": 50 * n / (t50 * 1e-3)})
    del X

    # adjoint: forward + backward of L = sum |qE - F_t|^2 + sum (phi - phi_t)^2, FP64
    s = workloads.adjoint()
    kw = dict(theta2=s.theta2, eps2=s.eps2, group_size=64, group_pack=64, precision="fp64")
    tgt = Solver(**kw)
    Xt = [torch.from_numpy(a).to(dev) for a in (s.x, s.y, s.z)]
    qt = torch.from_numpy(s.extra["q_target"]).to(dev)
    et, pt = coulomb_field(*Xt, qt, tgt)
    Ft, Pt = (qt[:, None] * et).detach(), pt.detach()
    tgt.close()
    sol = Solver(**kw)

    def fwd_bwd():
        Xg = [torch.from_numpy(a).to(dev).requires_grad_(True) for a in (s.x, s.y, s.z, s.q)]
        e, pot = coulomb_field(*Xg, sol)
        loss = ((Xg[3][:, None] * e - Ft) ** 2).sum() + ((pot - Pt) ** 2).sum()
        loss.backward()

    fwd_bwd()
    ms = _timed(fwd_bwd, 3)
    sol.close()
    recs.append({"workload": "adjoint_n262144", "n": s.n, "precision": "fp64", "group_size": 64,
                 "group_pack": 64, "forward_backward_ms": ms})
    return recs


def run_periodic(n, group, pack, precision, steps, warmup):
    """The periodic-boundary solve: a neutral +-1 plasma in the 3-D periodic
    unit cube (pepc-mini's charges), theta2 = 0.36, eps2 = 1e-4; each step is
    one grow + traverse = 27 mirror-box walks + the lattice far field."""
    import torch

    from pepc_amd import workloads
    from pepc_amd.solver import Solver
    s = workloads.periodic(n=n, seed=41)
    dev = torch.device("cuda", 0)
    x, y, z, q = (torch.from_numpy(a).to(dev) for a in (s.x, s.y, s.z, s.q))
    sol = Solver(theta2=s.theta2, eps2=s.eps2, group_size=group, precision=precision,
                 group_pack=pack)
    sol.set_periodic()
    out = None
    for _ in range(warmup):
        sol.grow(x, y, z, q, sync=False)
        out = sol.traverse(out, sync=False)
    sol.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sol.grow(x, y, z, q, sync=False)
        out = sol.traverse(out, sync=False)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    st = sol.sync()
    g, w = sol.timings()
    sol.close()
    return {"workload": f"periodic_plasma_n{n}_unit_cube", "n": n, "precision": precision,
            "group_size": group, "group_pack": pack, "ms_per_step": ms, "t_grow_ms": g, "t_walk_and_lattice_ms": w,
            "updates_per_s": n / (ms * 1e-3), "interactions": st["interactions"],
            "interactions_per_target": st["interactions"] / n}


def force_error(sol, out, sys_, nsamp):
    """Relative force and potential errors against the FP64 direct sum on
    nsamp sampled targets, and the relative error of the sampled potential
    energy sum_i q_i phi_i / 2 over the same targets (the tree's and the
    direct sum's)."""
    from pepc_amd import workloads
    keys, perm, xs, ys, zs, qs = sol.sorted_host()
    idx = workloads.sample_targets(len(xs), nsamp)
    e_ref, p_ref = sol.direct(xs, ys, zs, qs, idx)
    e = out.e[:, idx].T.cpu().numpy()
    er = e_ref.cpu().numpy()
    rel = np.linalg.norm(e - er, axis=1) / np.linalg.norm(er, axis=1)
    pot = out.pot[idx].cpu().numpy()
    pr = p_ref.cpu().numpy()
    prel = np.abs(pot - pr) / np.abs(pr)
    en_tree = 0.5 * float(np.sum(qs[idx] * pot))
    en_direct = 0.5 * float(np.sum(qs[idx] * pr))
    energy = {"targets": int(len(idx)), "tree": en_tree, "direct": en_direct,
              "rel_err": abs(en_tree - en_direct) / abs(en_direct)}
    return float(np.median(rel)), float(np.max(rel)), float(np.median(prel)), energy


def tree_error(out, ref):
    """Relative difference of two solves of the same grown tree (same order):
    the precision the FP32 + Kahan path loses against the FP64 tree at the
    same theta and expansion order, over every target."""
    e = (out.e - ref.e).norm(dim=0).cpu().numpy()
    en = ref.e.norm(dim=0).cpu().numpy()
    pd = (out.pot - ref.pot).abs().cpu().numpy()
    pn = ref.pot.abs().cpu().numpy()
    ok, okp = en > 0, pn > 0
    rel, prel = e[ok] / en[ok], pd[okp] / pn[okp]
    return {"targets": int(ok.sum()), "median_rel_force_err": float(np.median(rel)),
            "p999_rel_force_err": float(np.quantile(rel, 0.999)), "max_rel_force_err": float(rel.max()),
            "median_rel_pot_err": float(np.median(prel))}


def valu_issue(insts, t_walk_ms):
    """Wave64 VALU instructions per walk launch (SQ_INSTS_VALU, from the
    committed counter profile of the same workload) over this run's walk time,
    against the issue rate the FP32 peak implies."""
    if not insts:
        return None
    rate = insts / (t_walk_ms * 1e-3)
    return {"insts_per_launch": insts, "achieved_per_s": rate, "peak_per_s": PEAK_VALU_ISSUE,
            "frac": rate / PEAK_VALU_ISSUE,
            "source": "profiles/walk_explosion4m_pmc.txt (SQ_INSTS_VALU)"}


def run_theta_leg(theta, args, dev_inputs):
    """The same system at another opening angle: FP32 + Kahan and FP64 steps
    (grow + traverse) timed, force error vs the FP64 direct sum."""
    from pepc_amd import workloads
    s = workloads.explosion(n=args.n, theta=theta)
    steps = max(3, args.steps // 3)
    sol, out, st, ms, g, w = run_gpu(s, "fp32", args.group, steps, 1, dev_inputs, args.pack)
    med, mx, _, energy = force_error(sol, out, s, min(args.err_samples, 16384))
    sol.close()
    leg = {"theta": theta, "ms_per_step": ms, "updates_per_s": s.n / (ms * 1e-3),
           "t_grow_ms": g, "t_walk_ms": w, "interactions": st["interactions"],
           "median_rel_force_err": med, "max_rel_force_err": mx,
           "meets_error_bar": med <= ERROR_BAR, "sampled_energy_rel_err": energy["rel_err"],
           "err_samples": min(args.err_samples, 16384), "steps": steps}
    if not args.no_fp64:
        sol64, out64, st64, ms64, g64, w64 = run_gpu(s, "fp64", args.group, steps, 1, dev_inputs,
                                                      args.pack)
        sol64.close()
        leg["fp64_ms_per_step"] = ms64
        leg["speedup_fp32_over_fp64"] = ms64 / ms
    log(f"theta {theta}: fp32 {ms:.3f} ms/step, median err {med:.3e}"
        + (f", fp64 {leg['fp64_ms_per_step']:.3f} ms/step" if "fp64_ms_per_step" in leg else ""))
    return leg


def host_cpu():
    """Model name of this host's CPU (from /proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(sys_, group, stride, nthreads, pack=0, pp_stride=8):
    """The oracle (CPU restatement of the reference path, FP64) on this host:
    grow timed whole (keys on every host thread, presort, tree and upward pass
    serial as on one reference rank), then the group walk on every host thread
    at the reference's default tile rule (module_tree_walk_simple TILE_SIZE =
    16, no packing; the value), at the GPU run's rule, and per particle (the
    pthreads walk's one target per traversal, module_tree_walk_pthreads.f90)
    on every pp_stride-th particle, scaled."""
    from oracle import oracle as orc

    def grow():
        t0 = time.perf_counter()
        d = orc.domain_decompose(sys_.x, sys_.y, sys_.z, sys_.q)
        t = orc.tree_build(d.keys, d.x, d.y, d.z, d.q, d.box)
        return time.perf_counter() - t0, d, t

    grow()                       # warm-up
    t_grow, d, t = min((grow() for _ in range(3)), key=lambda r: r[0])

    def walk(G, P, k):
        tiles = orc.walk_tiles(d.keys, G, P)
        t2 = time.perf_counter()
        w = orc.walk_group(d.x, d.y, d.z, t, sys_.theta2, sys_.eps2, tile_size=G, sign=sys_.sign,
                           tile_stride=k, nthreads=nthreads, tiles=tiles)
        return (time.perf_counter() - t2) * k, len(tiles[0]), w.num_interactions

    walk(16, 0, stride * 8)      # warm-up (caches, page faults, thread pool)
    reps = [walk(16, 0, stride) for _ in range(3)]
    t_ref, nt_ref, ni_ref = min(reps)
    t_same, nt_same, ni_same = walk(group, pack, stride)
    t_pp, nt_pp, ni_pp = walk(1, 0, stride * pp_stride)
    return {"value": sys_.n / (t_grow + t_ref), "unit": "particle updates/s", "cores": nthreads,
            "kind": "port",
            "sample": (f"oracle (CPU restatement, FP64) on the same {sys_.n}-particle system: grow "
                       f"(box, keys on {nthreads} threads, reference MSD presort, tree, upward pass "
                       f"serial) timed whole ({t_grow:.2f} s); walk at the reference's default "
                       f"tile rule (G=16, {nt_ref} tiles, every {stride}-th, scaled) on "
                       f"{nthreads} threads ({t_ref:.2f} s); each phase one warm-up, then best "
                       f"of 3"),
            "t_grow_s": t_grow, "t_walk_s": t_ref, "interactions": ni_ref * stride,
            "same_rule_as_gpu": {"group_size": group, "group_pack": pack, "t_walk_s": t_same,
                                 "value": sys_.n / (t_grow + t_same), "tiles": nt_same},
            "per_particle": {"group_size": 1, "t_walk_s": t_pp, "value": sys_.n / (t_grow + t_pp),
                             "sample": f"every {stride * pp_stride}-th of {nt_pp} particles, "
                                       f"scaled", "interactions": ni_pp * stride * pp_stride},
            "host_cpu": host_cpu(), "host_threads_visible": os.cpu_count()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=4194304)
    ap.add_argument("--theta", type=float, default=0.6)
    ap.add_argument("--group", type=int, default=16)
    ap.add_argument("--pack", type=int, default=32,
                    help="merge runs of sibling walk tiles up to this many targets (0 = off)")
    ap.add_argument("--precision", default="fp32")
    ap.add_argument("--err-samples", type=int, default=65536)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-timestep", action="store_true",
                    help="skip the on-device leapfrog measurement (profiling runs)")
    ap.add_argument("--no-periodic", action="store_true",
                    help="skip the periodic-boundary measurement")
    ap.add_argument("--periodic-n", type=int, default=1048576)
    ap.add_argument("--periodic-group", type=int, default=64,
                    help="walk tile size of the periodic solve (each group walks 27 mirror "
                         "boxes, so larger groups amortise the traversal)")
    ap.add_argument("--periodic-pack", type=int, default=64)
    ap.add_argument("--periodic-only", action="store_true",
                    help="run only the periodic-boundary measurement (FP32, and FP64 unless "
                         "--no-fp64) and print its records")
    ap.add_argument("--other-configs", action="store_true",
                    help="time the other configurations (small FP64 system with its leapfrog "
                         "steps, 8M clustered gravity, 256-system ensemble with its graph "
                         "timestep, adjoint) and print one record each")
    ap.add_argument("--theta-legs", default="0.3",
                    help="comma-separated extra opening angles timed on the same system (FP32 "
                         "and FP64, fewer steps), empty for none")
    ap.add_argument("--cpu-stride", type=int, default=1)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run one plain device copy of MIB MiB (reference kernel for "
                         "scaling rocprofv3 FETCH_SIZE/WRITE_SIZE to bytes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU sharding "
              f"is a design note, not a measured path", file=sys.stderr)
        sys.exit(2)

    import torch

    if args.other_configs:
        for rec in run_other_configs(args.steps, args.warmup):
            print(json.dumps(rec), flush=True)
        return

    if args.periodic_only:
        for prec in ("fp32",) if args.no_fp64 else ("fp32", "fp64"):
            per = run_periodic(args.periodic_n, args.periodic_group, args.periodic_pack, prec, args.steps,
                               args.warmup)
            print(json.dumps(per), flush=True)
        return

    from pepc_amd import workloads
    s = workloads.explosion(n=args.n, theta=args.theta)
    dev = torch.device("cuda", 0)
    if args.calibrate_copy:
        a = torch.ones(args.calibrate_copy * (1 << 20) // 4, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        log(f"calibration copy: {a.numel() * 4} bytes read, {a.numel() * 4} bytes written")
        del a, b
    dev_inputs = [torch.from_numpy(a).to(dev) for a in (s.x, s.y, s.z, s.q)]
    log(f"workload explosion n={s.n} theta={args.theta} G={args.group} {args.precision}")

    sol, out, st, ms, t_grow_ms, t_walk_ms = run_gpu(s, args.precision, args.group, args.steps,
                                                     args.warmup, dev_inputs, args.pack)
    log(f"{args.precision}: {ms:.3f} ms/step (grow {t_grow_ms:.3f} ms, walk {t_walk_ms:.3f} ms) "
        f"stats {json.dumps({k: v for k, v in st.items() if k != 'level_offset'})}")
    med, mx, pmed, energy = force_error(sol, out, s, args.err_samples)
    log(f"force error vs FP64 direct sum ({args.err_samples} targets): median {med:.3e} max {mx:.3e}; "
        f"sampled energy rel err {energy['rel_err']:.3e}")
    flops = walk_flops(st)
    achieved = flops / (t_walk_ms * 1e-3) / 1e12
    traffic = valu_insts = None
    tp = os.path.join(ROOT, "profiles", "walk_explosion4m_traffic.json")
    if os.path.exists(tp):
        try:
            rec = json.load(open(tp))
            traffic = rec.get("hbm_bytes_per_launch")
            valu_insts = rec.get("sq_insts_valu_per_launch")
        except Exception:
            traffic = valu_insts = None
    sol.close()

    fp64 = None
    if not args.no_fp64 and args.precision != "fp64":
        sol64, out64, st64, ms64, g64, w64 = run_gpu(s, "fp64", args.group, max(2, args.steps // 2),
                                                      1, dev_inputs, args.pack)
        med64, mx64, _, _ = force_error(sol64, out64, s, min(args.err_samples, 16384))
        fp64 = {"ms_per_step": ms64, "t_walk_ms": w64, "t_grow_ms": g64,
                "median_rel_force_err": med64, "speedup_fp32_over_fp64": ms64 / ms,
                "fp32_vs_fp64_tree": tree_error(out, out64)}
        fp64["fp32_vs_fp64_tree"]["meets_error_bar"] = (
            fp64["fp32_vs_fp64_tree"]["median_rel_force_err"] <= ERROR_BAR)
        log(f"fp64: {ms64:.3f} ms/step walk {w64:.3f} ms; median err {med64:.3e}; "
            f"fp32 speedup {ms64 / ms:.2f}x; fp32 vs fp64 tree {json.dumps(fp64['fp32_vs_fp64_tree'])}")
        sol64.close()

    legs = {}
    for th in [float(v) for v in args.theta_legs.split(",") if v.strip()]:
        legs[f"theta_{th:g}"] = run_theta_leg(th, args, dev_inputs)

    ts_ms = ts_drift = None
    if not args.no_timestep:
        ts_ms, ts_drift = run_timestep(s, args.group, args.pack, max(args.steps, 5))
        log(f"on-device leapfrog step (graph): {ts_ms:.3f} ms/step, energy drift {ts_drift:.2e}")

    per = None
    if not args.no_periodic:
        per = run_periodic(args.periodic_n, args.periodic_group, args.periodic_pack, args.precision,
                           max(3, args.steps // 2), 1)
        log(f"periodic: {per['ms_per_step']:.3f} ms/step (grow {per['t_grow_ms']:.3f} ms, "
            f"walk + lattice {per['t_walk_and_lattice_ms']:.3f} ms) "
            f"{per['interactions_per_target']:.0f} interactions/target")

    cpu = None
    if not args.no_cpu:
        nthr = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = cpu_baseline(s, args.group, args.cpu_stride, nthr, args.pack)
        log(f"cpu baseline: {cpu['value']:.3e} updates/s ({cpu['sample']})")

    rec = {
        "metric": "Coulomb particle updates/s, N=4M \u03b8=0.6, FP32c vs FP64 at rel force err \u22641e-4",
        "value": s.n / (ms * 1e-3),
        "unit": "particle updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+kahan (f64 upward pass and MAC)" if args.precision == "fp32" else "f64",
        "data": "synthetic (seeded Coulomb explosion, uniform unit ball)",
        "config": {"workload": f"coulomb_explosion_n{s.n}_theta{args.theta}", "n": s.n,
                   "theta": args.theta, "eps": 1e-3, "expansion": "quadrupole",
                   "group_size": args.group, "group_pack": args.pack,
                   "precision": args.precision,
                   "t_grow_ms": t_grow_ms, "t_walk_ms": t_walk_ms,
                   "interactions_per_s": st["interactions"] / (t_walk_ms * 1e-3),
                   "median_rel_force_err": med, "max_rel_force_err": mx,
                   "median_rel_pot_err": pmed, "err_samples": args.err_samples,
                   "meets_error_bar": med <= ERROR_BAR,
                   "error_bar_note": "median relative force error vs the FP64 direct sum <= 1e-4; "
                                     "at theta 0.6 the Barnes-Hut truncation error (the FP64 tree's "
                                     "as well) is above it, the FP32 + Kahan vs FP64 tree "
                                     "difference (fp64_path.fp32_vs_fp64_tree) is below it, and "
                                     "the theta_0.3 leg meets it",
                   "sampled_energy": energy,
                   "fp64_path": fp64,
                   "theta_legs": legs,
                   "timestep_ms_per_step": ts_ms,
                   "timestep_note": "kick-drift-kick step with grow from the previous step's key "
                                    "order, HIP-graph replay (dt=1e-3)",
                   "timestep_energy_drift": ts_drift,
                   "periodic": per},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_VALU_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_VALU_TFLOPS,
                     "traffic": traffic, "kernel": "walk_kernel<float>",
                     "flops_per_launch": flops,
                     "flops_note": "reference formula count 22 N_P2P + 135 N_M2P + 9 N_MAC",
                     "valu_issue": valu_issue(valu_insts, t_walk_ms),
                     "note": "bound by FP32 arithmetic issued on the vector ALUs, neither HBM "
                             "(traffic well below 8 TB/s x the launch) nor matrix cores (no dense "
                             "contraction); peak = MI355X FP32 vector 157.3 TFLOP/s"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
