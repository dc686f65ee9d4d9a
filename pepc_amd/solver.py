"""Python front end of the GPU tree solve (thin layer over the C-ABI).

``Solver`` owns one ``pepc_gpu`` context.  Device memory for inputs and
outputs comes from torch (plumbing only); all computation happens in
``libpepc_gpu.so`` on torch's current stream.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib


def _ptr(t):
    return C.c_void_p(t.data_ptr())


@dataclass
class TraverseResult:
    e: "object"        # (n, 3) float64 tensor, sorted order
    pot: "object"
    work: "object"


class Solver:
    """One GPU context: ``grow`` then ``traverse``, as pepc_grow_tree and
    pepc_traverse_tree (module_pepc.f90:435-468)."""

    def __init__(self, theta2=0.36, eps2=0.0, group_size=16, precision="fp32",
                 sign=1.0, node_factor=2.5, force_cubic_domain=False, walk_blocks=0,
                 device=0, group_pack=0, list_tile=0, mac_select=0):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("pepc_amd.Solver needs a GPU (no HIP device visible)")
        self.torch = torch
        self.lib = _lib.load()
        prec = {"fp32": _lib.FP32_KAHAN, "fp64": _lib.FP64}[precision] \
            if isinstance(precision, str) else int(precision)
        self.cfg = _lib.default_config(theta2=float(theta2), eps2=float(eps2),
                                       group_size=int(group_size), precision=prec,
                                       sign=float(sign), node_factor=float(node_factor),
                                       force_cubic_domain=int(bool(force_cubic_domain)),
                                       walk_blocks=int(walk_blocks), device=int(device),
                                       group_pack=int(group_pack), list_tile=int(list_tile),
                                       mac_select=int(mac_select))
        self.device = torch.device("cuda", device)
        h = C.c_void_p()
        rc = self.lib.pepc_gpu_create(C.byref(h), C.byref(self.cfg))
        if rc:
            raise _lib.PepcGpuError(rc, "pepc_gpu_create failed")
        self.h = h
        self.n = 0
        self._bind_stream()

    def _bind_stream(self):
        s = self.torch.cuda.current_stream(self.device)
        self._check(self.lib.pepc_gpu_set_stream(self.h, C.c_void_p(s.cuda_stream)))

    def _check(self, rc):
        if rc:
            buf = C.create_string_buffer(512)
            self.lib.pepc_gpu_last_error(self.h, buf, 512)
            raise _lib.PepcGpuError(rc, buf.value.decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.pepc_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------ core
    def _dev(self, a):
        t = self.torch
        if isinstance(a, np.ndarray):
            return t.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(self.device)
        return a.to(device=self.device, dtype=t.float64).contiguous()

    def _grow_synced(self, launch):
        """Runs a grow launch and waits for it.  A node pool too small for the
        tree (the build counts its nodes before it stops) is enlarged to the
        counted size and the grow runs again, as pepc_gpu_grow does."""
        launch()
        st = _lib.Stats()
        rc = self.lib.pepc_gpu_sync(self.h, C.byref(st))
        if rc == _lib.EOVERFLOW:
            need = int(st.nnodes)
            self._check(self.lib.pepc_gpu_reserve_nodes(self.h, need + need // 16 + 64))
            launch()
            rc = self.lib.pepc_gpu_sync(self.h, C.byref(st))
        self._check(rc)
        return st.as_dict()

    def grow(self, x, y, z, q, sync=True):
        """Builds the tree.  Inputs: numpy arrays or device tensors (float64).
        With sync, a node pool overflow enlarges the pool and grows again;
        without, it is reported by the next sync."""
        self._bind_stream()
        xs = [self._dev(v) for v in (x, y, z, q)]
        self._inputs = xs      # keep alive until the stream has consumed them
        n = xs[0].numel()
        launch = lambda: self._check(self.lib.pepc_gpu_grow_device(self.h, n, *[_ptr(v) for v in xs]))
        self.n = n
        if sync:
            return self._grow_synced(launch)
        launch()
        return None

    def grow_batch(self, x, y, z, q, offsets, sync=True):
        """Builds one tree per system: system s holds particles
        [offsets[s], offsets[s+1]) of the concatenated inputs."""
        self._bind_stream()
        xs = [self._dev(v) for v in (x, y, z, q)]
        self._inputs = xs
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        nsys = len(off) - 1
        if nsys < 1 or off[-1] != xs[0].numel():
            raise ValueError("grow_batch: offsets must have nsys + 1 entries ending at n")
        launch = lambda: self._check(self.lib.pepc_gpu_grow_batch_device(
            self.h, nsys, off.ctypes.data_as(C.c_void_p), *[_ptr(v) for v in xs]))
        self.n = int(off[-1])
        self.offsets = off
        if sync:
            return self._grow_synced(launch)
        launch()
        return None

    def traverse(self, out=None, sync=True):
        """Fields, potentials and work for the grown particles (sorted order)."""
        t = self.torch
        n = self.n
        if out is None:
            out = TraverseResult(t.empty((3, n), dtype=t.float64, device=self.device),
                                 t.empty(n, dtype=t.float64, device=self.device),
                                 t.empty(n, dtype=t.float64, device=self.device))
        e = out.e
        self._check(self.lib.pepc_gpu_traverse_device(self.h, _ptr(e[0]), _ptr(e[1]), _ptr(e[2]),
                                                      _ptr(out.pot), _ptr(out.work)))
        if sync:
            self.sync()
        return out

    def traverse_sinks(self, x, y, z, grouped=False, sync=True):
        """Fields, potentials and interaction counts (work) at other points
        (any order; results in the same order) from the grown tree, as the
        reference's traversal of another particle array."""
        t = self.torch
        xs = [self._dev(v) for v in (x, y, z)]
        self._sinks = xs
        m = xs[0].numel()
        out = TraverseResult(t.empty((3, m), dtype=t.float64, device=self.device),
                             t.empty(m, dtype=t.float64, device=self.device),
                             t.empty(m, dtype=t.float64, device=self.device))
        self._bind_stream()
        mode = _lib.SINKS_GROUPED if grouped else _lib.SINKS_PER_PARTICLE
        self._check(self.lib.pepc_gpu_traverse_sinks_device(
            self.h, m, *[_ptr(v) for v in xs], mode, _ptr(out.e[0]), _ptr(out.e[1]), _ptr(out.e[2]),
            _ptr(out.pot), _ptr(out.work)))
        if sync:
            self.sync()
        return out

    def sink_counts(self, m):
        """(p2p, m2p) per sink of the last traverse_sinks (sinks' order)."""
        p2p = np.empty(m, np.uint32)
        m2p = np.empty(m, np.uint32)
        vp = lambda a: a.ctypes.data_as(C.c_void_p)
        self._check(self.lib.pepc_gpu_read_sink_counts(self.h, m, vp(p2p), vp(m2p)))
        return p2p, m2p

    def check_tree(self) -> int:
        """Number of consistency violations of the grown tree (0 = sound)."""
        nb = C.c_int64()
        self._check(self.lib.pepc_gpu_check_tree(self.h, C.byref(nb)))
        return nb.value

    def sync(self) -> dict:
        st = _lib.Stats()
        self._check(self.lib.pepc_gpu_sync(self.h, C.byref(st)))
        return st.as_dict()

    def timings(self):
        a, b = C.c_double(), C.c_double()
        self._check(self.lib.pepc_gpu_timings(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def autotune(self, x, y, z, q, offsets=None, reps=2):
        """Times every candidate configuration on these particles and keeps the
        fastest (see pepc_gpu_autotune).  Returns (best, candidates) as dicts."""
        self._bind_stream()
        xs = [self._dev(v) for v in (x, y, z, q)]
        n = xs[0].numel()
        off = None if offsets is None else np.ascontiguousarray(offsets, dtype=np.int64)
        nsys = 1 if off is None else len(off) - 1
        if off is not None and off[-1] != n:
            raise ValueError("autotune: offsets must end at n")
        cap = 64
        out = (_lib.TuneEntry * cap)()
        nc = C.c_int64()
        self._check(self.lib.pepc_gpu_autotune(
            self.h, n, nsys, None if off is None else off.ctypes.data_as(C.c_void_p),
            *[_ptr(v) for v in xs], int(reps), out, cap, C.byref(nc)))
        self._check(self.lib.pepc_gpu_get_config(self.h, C.byref(self.cfg)))
        k = min(nc.value + 1, cap)
        rows = [out[i].as_dict() for i in range(k)]
        return rows[0], rows[1:]

    def config(self) -> dict:
        cfg = _lib.Config()
        self._check(self.lib.pepc_gpu_get_config(self.h, C.byref(cfg)))
        return {k: getattr(cfg, k) for k, _ in cfg._fields_}

    def freeze(self) -> int:
        """Records the interaction lists of the last grow; returns their size."""
        self._bind_stream()
        ne = C.c_int64()
        self._check(self.lib.pepc_gpu_freeze(self.h, C.byref(ne)))
        return ne.value

    def replay(self, x, y, z, q, out=None, sync=True):
        """FP64 fields of the frozen lists for new particle data (input order
        of the last grow); sorted order out, as traverse."""
        t = self.torch
        n = self.n
        xs = [self._dev(v) for v in (x, y, z, q)]
        if any(v.numel() != n for v in xs):
            raise ValueError("replay: arrays must have the grown size")
        if out is None:
            out = TraverseResult(t.empty((3, n), dtype=t.float64, device=self.device),
                                 t.empty(n, dtype=t.float64, device=self.device), None)
        self._bind_stream()
        e = out.e
        self._check(self.lib.pepc_gpu_replay_device(self.h, *[_ptr(v) for v in xs], _ptr(e[0]),
                                                    _ptr(e[1]), _ptr(e[2]), _ptr(out.pot)))
        self._replay_inputs = xs
        if sync:
            self.sync()
        return out

    def adjoint(self, ge, gpot, sync=True):
        """dL/dx, dL/dy, dL/dz, dL/dq (sorted order, float64 device tensors)
        for upstream gradients ge = dL/dE (3, n) and gpot = dL/dpot (n) of the
        last grow, sorted order; tree and interaction lists held fixed."""
        t = self.torch
        n = self.n
        ge = self._dev(ge)
        gp = self._dev(gpot)
        if ge.shape != (3, n) or gp.shape != (n,):
            raise ValueError("adjoint: ge must be (3, n) and gpot (n,)")
        out = [t.empty(n, dtype=t.float64, device=self.device) for _ in range(4)]
        self._bind_stream()
        self._check(self.lib.pepc_gpu_adjoint_device(self.h, _ptr(ge[0]), _ptr(ge[1]), _ptr(ge[2]),
                                                     _ptr(gp), *[_ptr(o) for o in out]))
        self._keep = (ge, gp)
        if sync:
            self.sync()
        return out

    def restore(self, *arrays):
        """Device arrays in sorted order (e.g. traverse results) back to the
        input order of the last grow; returns new tensors."""
        t = self.torch
        ins = [self._dev(a).contiguous() for a in arrays]
        if any(a.numel() != self.n for a in ins):
            raise ValueError("restore: arrays must have the grown size")
        outs = [t.empty_like(a) for a in ins]
        k = len(ins)
        pin = (C.c_void_p * k)(*[a.data_ptr() for a in ins])
        pout = (C.c_void_p * k)(*[a.data_ptr() for a in outs])
        self._bind_stream()
        self._check(self.lib.pepc_gpu_restore_device(self.h, k, pin, pout))
        self._keep_restore = ins
        t.cuda.synchronize(self.device)
        return outs

    # -------------------------------------------------------------- timestep
    def state_set(self, x, y, z, v, q, m, offsets=None, max_records=1024):
        """Loads particle state (host arrays, input order) for the on-device
        kick-drift-kick timestep and runs the initial solve."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in
                (x, y, z, v[:, 0], v[:, 1], v[:, 2], q, m)]
        n = len(arrs[0])
        if offsets is None:
            off = np.array([0, n], np.int64)
        else:
            off = np.ascontiguousarray(offsets, dtype=np.int64)
        self._bind_stream()
        vp = lambda a: a.ctypes.data_as(C.c_void_p)
        self._check(self.lib.pepc_gpu_state_set(self.h, n, len(off) - 1, vp(off), *[vp(a) for a in arrs],
                                                int(max_records)))
        self.n = n
        self.offsets = off
        self._state_arrays = arrs
        self._max_records = int(max_records)

    def run(self, nsteps, dt, graph=False, sync=True):
        self._bind_stream()
        self._check(self.lib.pepc_gpu_run(self.h, int(nsteps), float(dt),
                                          _lib.RUN_GRAPH if graph else 0))
        if sync:
            return self.sync()
        return None

    def state_get(self):
        """Host copies of the state: dict of x, y, z, v (n, 3), q, m, label."""
        n = self.n
        a = [np.empty(n) for _ in range(8)]
        lab = np.empty(n, np.int64)
        vp = lambda t: t.ctypes.data_as(C.c_void_p)
        self._check(self.lib.pepc_gpu_state_get(self.h, *[vp(t) for t in a], vp(lab)))
        return {"x": a[0], "y": a[1], "z": a[2], "v": np.stack(a[3:6], 1), "q": a[6], "m": a[7],
                "label": lab}

    def energies(self):
        """(records, nsys, 2) array of {1/2 sum q phi, 1/2 sum m v^2} per solve."""
        nr = C.c_int64()
        self._check(self.lib.pepc_gpu_read_energies(self.h, 0, None, C.byref(nr)))
        nsys = len(self.offsets) - 1
        k = min(nr.value, self._max_records)
        out = np.empty((k, nsys, 2))
        self._check(self.lib.pepc_gpu_read_energies(self.h, k, out.ctypes.data_as(C.c_void_p),
                                                    C.byref(nr)))
        return out

    # ------------------------------------------------------------- read-back
    def sorted_host(self):
        """(keys uint64, perm int64, x, y, z, q) of the last grow, on the host."""
        n = self.n
        keys = np.empty(n, np.uint64)
        perm = np.empty(n, np.int64)
        arrs = [np.empty(n) for _ in range(4)]
        vp = lambda a: a.ctypes.data_as(C.c_void_p)
        self._check(self.lib.pepc_gpu_read_sorted(self.h, vp(keys), vp(perm), *[vp(a) for a in arrs]))
        return (keys, perm, *arrs)

    def read_tree(self):
        st = self.sync()
        m = st["nnodes"]
        link = np.empty((m, 4), np.uint32)
        first = np.empty(m, np.uint32)
        mom = np.empty((m, 16))
        nn = C.c_int64()
        self._check(self.lib.pepc_gpu_read_tree(self.h, m, link.ctypes.data_as(C.c_void_p),
                                                first.ctypes.data_as(C.c_void_p),
                                                mom.ctypes.data_as(C.c_void_p), C.byref(nn)))
        return {"link": link, "first": first, "mom": mom, "stats": st}

    def read_walk(self):
        st = self.sync()
        g = st["ngroups"]
        gf = np.empty(g, np.uint32)
        gc = np.empty(g, np.uint32)
        gm = np.empty(g, np.uint32)
        p2p = np.empty(self.n, np.uint32)
        m2p = np.empty(self.n, np.uint32)
        ng = C.c_int64()
        vp = lambda a: a.ctypes.data_as(C.c_void_p)
        self._check(self.lib.pepc_gpu_read_walk(self.h, g, vp(gf), vp(gc), vp(gm), vp(p2p), vp(m2p),
                                                C.byref(ng)))
        return {"gfirst": gf, "gcount": gc, "gmac": gm, "p2p": p2p, "m2p": m2p, "stats": st}

    # ------------------------------------------------------ periodic boundaries
    def set_periodic(self, periodicity=(1, 1, 1), lattice=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0),
                     layers=1, far_field=True, correction=_lib.EXTRINSIC_FICTCHARGE):
        """Periodic boundaries for the following traverses (periodicity=None
        clears them).  lattice: three box lengths (principal axes) or a 3x3
        array whose rows are t_lattice_1..3; correction: the far field's
        extrinsic-to-intrinsic correction (fmm_extrinsic_correction)."""
        if periodicity is None:
            self._check(self.lib.pepc_gpu_set_periodic(self.h, None))
            return
        p = _lib.Periodic()
        self.lib.pepc_gpu_periodic_default(C.byref(p))
        lat = np.asarray(lattice, dtype=np.float64)
        if lat.shape == (3,):
            lat = np.diag(lat)
        if lat.shape != (3, 3):
            raise ValueError("set_periodic: lattice must be 3 lengths or a 3x3 array")
        for k in range(3):
            p.periodicity[k] = int(bool(periodicity[k]))
            p.origin[k] = float(origin[k])
        for k, v in enumerate(lat.reshape(9)):
            p.t_lattice[k] = float(v)
        p.mirror_box_layers = int(layers)
        p.far_field = int(bool(far_field))
        p.extrinsic_correction = int(correction)
        self._check(self.lib.pepc_gpu_set_periodic(self.h, C.byref(p)))

    def lattice(self):
        """(ml complex (861,), mu complex (231,), fict (nfict, 4)) of the far field."""
        ml = np.empty(2 * 861)
        mu = np.empty(2 * 231)
        fict = np.empty(16)
        nf = C.c_int32()
        vp = lambda a: a.ctypes.data_as(C.c_void_p)
        self._check(self.lib.pepc_gpu_lattice_get(self.h, vp(ml), vp(mu), vp(fict), C.byref(nf)))
        return ml.view(np.complex128), mu.view(np.complex128), fict[:4 * nf.value].reshape(-1, 4)

    def constrain_periodic(self, x, y, z):
        """Wraps device tensors x, y, z (float64, in place) into the central box."""
        self._bind_stream()
        n = x.numel()
        self._check(self.lib.pepc_gpu_constrain_periodic_device(self.h, n, _ptr(x), _ptr(y), _ptr(z)))
        self.torch.cuda.synchronize(self.device)

    def direct(self, x, y, z, q, testidx, eps2=None, sign=None):
        """FP64 direct sum at the given targets (device computation)."""
        t = self.torch
        xs = [self._dev(v) for v in (x, y, z, q)]
        idx = np.asarray(testidx, dtype=np.int64)
        if idx.size and (idx.min() < 0 or idx.max() >= xs[0].numel()):
            raise ValueError("direct: target index out of range")
        ti = t.as_tensor(idx).to(self.device)
        m = ti.numel()
        out = [t.empty(m, dtype=t.float64, device=self.device) for _ in range(4)]
        self._bind_stream()
        self._check(self.lib.pepc_gpu_direct_device(
            self.h, xs[0].numel(), *[_ptr(v) for v in xs], m, _ptr(ti),
            float(self.cfg.eps2 if eps2 is None else eps2),
            float(self.cfg.sign if sign is None else sign), *[_ptr(o) for o in out]))
        t.cuda.synchronize(self.device)
        return t.stack(out[:3], 1), out[3]

