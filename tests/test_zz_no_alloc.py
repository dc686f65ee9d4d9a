"""pepc_gpu_traverse_sinks_noalloc (static device scratch, no allocation)
against pepc_gpu_traverse_sinks_device, through the C-ABI.

Both entry points run the same kernels on the same grown tree, so fields,
potentials, work and per-sink counts must be identical bit for bit, in both
sink modes, on two calls in a row (the second with fewer sinks, so stale
scratch contents from the first call would show), and the context's own
sink path must still agree after the scratch path ran.  Sink counts above
the scratch capacity are refused before any device work.
"""
import ctypes as C

import numpy as np
import pytest

from pepc_amd import _lib, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def grown():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pepc_amd.solver import Solver
    s = workloads.mini(n=20000, seed=61)
    sol = Solver(theta2=s.theta2, eps2=s.eps2, precision="fp32")
    sol.grow(s.x, s.y, s.z, s.q)
    yield sol, s
    sol.close()


def _sinks(m, seed, s):
    r = np.random.default_rng(seed)
    p = r.uniform(-0.1, 1.1, (m, 3))
    k = min(m // 10, s.n)
    pick = r.choice(s.n, k, replace=False)
    p[:k, 0], p[:k, 1], p[:k, 2] = s.x[pick], s.y[pick], s.z[pick]   # some on the sources
    return p


def _outputs(torch, m):
    return [torch.full((m,), np.nan, dtype=torch.float64, device="cuda") for _ in range(5)]


def _call(sol, fn, pts, mode, outs):
    ptr = lambda t: C.c_void_p(t.data_ptr())
    rc = fn(sol.h, pts[0].numel(), *[ptr(t) for t in pts], mode, *[ptr(t) for t in outs])
    torch = sol.torch
    torch.cuda.synchronize()
    return rc


def _counts(sol, m):
    p2p = np.empty(m, np.uint32)
    m2p = np.empty(m, np.uint32)
    rc = sol.lib.pepc_gpu_read_sink_counts(sol.h, m, p2p.ctypes.data_as(C.c_void_p),
                                           m2p.ctypes.data_as(C.c_void_p))
    assert rc == _lib.OK
    return p2p, m2p


@pytest.mark.parametrize("mode", [_lib.SINKS_PER_PARTICLE, _lib.SINKS_GROUPED])
def test_noalloc_equals_device_entry(grown, mode):
    import torch
    sol, s = grown
    lib = sol.lib
    for m, seed in ((150000, 71), (3000, 72)):   # second call smaller than the first
        print(f"noalloc sinks: mode {mode} m {m}", flush=True)
        p = _sinks(m, seed, s)
        pts = [torch.from_numpy(p[:, d].copy()).to("cuda") for d in range(3)]
        ref = _outputs(torch, m)
        assert _call(sol, lib.pepc_gpu_traverse_sinks_device, pts, mode, ref) == _lib.OK
        ref_counts = _counts(sol, m)
        got = _outputs(torch, m)
        assert _call(sol, lib.pepc_gpu_traverse_sinks_noalloc, pts, mode, got) == _lib.OK
        got_counts = _counts(sol, m)
        for a, b in zip(got, ref):
            assert torch.equal(a, b)
        assert np.array_equal(got_counts[0], ref_counts[0])
        assert np.array_equal(got_counts[1], ref_counts[1])
        assert int(ref_counts[0].sum() + ref_counts[1].sum()) > 0
        # the context's own path still agrees after the scratch path ran
        again = _outputs(torch, m)
        assert _call(sol, lib.pepc_gpu_traverse_sinks_device, pts, mode, again) == _lib.OK
        for a, b in zip(again, ref):
            assert torch.equal(a, b)


def test_noalloc_refuses_more_sinks_than_its_scratch(grown):
    import torch
    sol, s = grown
    m = _lib.NOALLOC_SINKS + 1
    pts = [torch.zeros(m, dtype=torch.float64, device="cuda") for _ in range(3)]
    outs = _outputs(torch, m)
    rc = _call(sol, sol.lib.pepc_gpu_traverse_sinks_noalloc, pts, _lib.SINKS_GROUPED, outs)
    assert rc == _lib.EINVAL
    buf = C.create_string_buffer(256)
    sol.lib.pepc_gpu_last_error(sol.h, buf, 256)
    assert b"262144" in buf.value
    assert torch.isnan(outs[0]).all()   # nothing ran
