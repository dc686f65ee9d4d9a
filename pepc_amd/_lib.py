"""ctypes binding of ``libpepc_gpu.so`` (the C-ABI of ``include/pepc_gpu.h``).

The product path: every call goes to the HIP library.  There is no CPU
fallback; importing works without a GPU (so the exports can be checked), but
creating a context raises if the library is missing or no device is present.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# PEPC_GPU_LIB selects another build of the same C-ABI (e.g. a tuning variant)
LIB_PATH = os.environ.get("PEPC_GPU_LIB") or os.path.join(HERE, "libpepc_gpu.so")

FP32_KAHAN = 0
FP64 = 1

OK, EINVAL, EHIP, ENOMEM, EOVERFLOW, EKEYS, EBOX, EWALK = 0, -1, -2, -3, -4, -5, -6, -7

# Entry points declared in include/pepc_gpu.h (checked by tests/test_abi.py).
EXPORTS = [
    "pepc_gpu_config_default", "pepc_gpu_create", "pepc_gpu_destroy", "pepc_gpu_last_error",
    "pepc_gpu_set_stream", "pepc_gpu_grow", "pepc_gpu_traverse", "pepc_gpu_grow_device",
    "pepc_gpu_traverse_device", "pepc_gpu_sync", "pepc_gpu_timber", "pepc_gpu_sorted_device",
    "pepc_gpu_read_sorted", "pepc_gpu_read_tree", "pepc_gpu_read_walk", "pepc_gpu_direct_device",
    "pepc_gpu_timings", "pepc_gpu_grow_batch_device", "pepc_gpu_state_set", "pepc_gpu_run",
    "pepc_gpu_state_get", "pepc_gpu_state_device", "pepc_gpu_read_energies",
    "pepc_gpu_adjoint_device", "pepc_gpu_autotune", "pepc_gpu_get_config", "pepc_gpu_freeze",
    "pepc_gpu_replay_device", "pepc_gpu_restore_device", "pepc_gpu_periodic_default",
    "pepc_gpu_set_periodic", "pepc_gpu_lattice_get", "pepc_gpu_constrain_periodic_device",
    "pepc_gpu_traverse_sinks_device", "pepc_gpu_traverse_sinks", "pepc_gpu_read_sink_counts",
    "pepc_gpu_check_tree", "pepc_gpu_direct", "pepc_gpu_reserve_nodes", "pepc_gpu_read_box",
    "pepc_gpu_traverse_sinks_noalloc",
]

# Sink capacity of pepc_gpu_traverse_sinks_noalloc (its static device scratch).
NOALLOC_SINKS = 262144

SINKS_PER_PARTICLE = 0
SINKS_GROUPED = 1

RUN_GRAPH = 1


class Config(C.Structure):
    _fields_ = [("theta2", C.c_double), ("eps2", C.c_double), ("sign", C.c_double),
                ("node_factor", C.c_double), ("mac_select", C.c_int32),
                ("group_size", C.c_int32), ("precision", C.c_int32),
                ("force_cubic_domain", C.c_int32), ("walk_blocks", C.c_int32),
                ("device", C.c_int32), ("group_pack", C.c_int32), ("list_tile", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("n", C.c_int64), ("nnodes", C.c_int64), ("nleaf", C.c_int64),
                ("ntwig", C.c_int64), ("ngroups", C.c_int64), ("dup_fixed", C.c_int64),
                ("interactions", C.c_uint64), ("mac_evaluations", C.c_uint64),
                ("p2p", C.c_uint64), ("m2p", C.c_uint64), ("amb_nodes", C.c_uint64),
                ("walk_steps", C.c_uint64), ("flags", C.c_uint32),
                ("maxlevel", C.c_int32), ("level_offset", C.c_uint32 * 22)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "level_offset"}
        d["level_offset"] = list(self.level_offset)
        return d


class TuneEntry(C.Structure):
    _fields_ = [("group_size", C.c_int32), ("group_pack", C.c_int32), ("list_tile", C.c_int32),
                ("walk_blocks", C.c_int32), ("t_grow_ms", C.c_double), ("t_walk_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Periodic(C.Structure):
    _fields_ = [("periodicity", C.c_int32 * 3), ("mirror_box_layers", C.c_int32),
                ("t_lattice", C.c_double * 9), ("origin", C.c_double * 3),
                ("far_field", C.c_int32), ("extrinsic_correction", C.c_int32)]


# fmm_extrinsic_correction values (pepc_gpu_periodic.extrinsic_correction)
EXTRINSIC_NONE, EXTRINSIC_REDLACK, EXTRINSIC_FICTCHARGE = 0, 1, 2


class PepcGpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pepc_gpu error {code}: {msg}")
        self.code = code


_lib = None

_vp = C.c_void_p
_dp = C.POINTER(C.c_double)
_i64 = C.c_int64


def load(path: str = LIB_PATH):
    """Load and prototype the library; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libpepc_gpu.so not built ({path}); run python -m pepc_amd.build")
    L = C.CDLL(path)
    L.pepc_gpu_config_default.argtypes = [C.POINTER(Config)]
    L.pepc_gpu_config_default.restype = None
    L.pepc_gpu_create.argtypes = [C.POINTER(_vp), C.POINTER(Config)]
    L.pepc_gpu_destroy.argtypes = [_vp]
    L.pepc_gpu_last_error.argtypes = [_vp, C.c_char_p, C.c_size_t]
    L.pepc_gpu_set_stream.argtypes = [_vp, _vp]
    L.pepc_gpu_grow.argtypes = [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    L.pepc_gpu_traverse.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp]
    L.pepc_gpu_grow_device.argtypes = [_vp, _i64, _vp, _vp, _vp, _vp]
    L.pepc_gpu_traverse_device.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp]
    L.pepc_gpu_sync.argtypes = [_vp, C.POINTER(Stats)]
    L.pepc_gpu_timber.argtypes = [_vp]
    L.pepc_gpu_sorted_device.argtypes = [_vp] + [C.POINTER(_vp)] * 6
    L.pepc_gpu_read_sorted.argtypes = [_vp] * 7
    L.pepc_gpu_read_tree.argtypes = [_vp, _i64, _vp, _vp, _vp, C.POINTER(_i64)]
    L.pepc_gpu_read_walk.argtypes = [_vp, _i64, _vp, _vp, _vp, _vp, _vp, C.POINTER(_i64)]
    L.pepc_gpu_direct_device.argtypes = [_vp, _i64, _vp, _vp, _vp, _vp, _i64, _vp,
                                         C.c_double, C.c_double, _vp, _vp, _vp, _vp]
    L.pepc_gpu_direct.argtypes = [_vp, _i64, _vp, _vp, _vp, _vp, _i64, _vp,
                                  C.c_double, C.c_double, _vp, _vp, _vp, _vp]
    L.pepc_gpu_timings.argtypes = [_vp, _dp, _dp]
    L.pepc_gpu_grow_batch_device.argtypes = [_vp, _i64, _vp, _vp, _vp, _vp, _vp]
    L.pepc_gpu_state_set.argtypes = [_vp, _i64, _i64, _vp] + [_vp] * 8 + [_i64]
    L.pepc_gpu_run.argtypes = [_vp, _i64, C.c_double, C.c_int]
    L.pepc_gpu_state_get.argtypes = [_vp] * 10
    L.pepc_gpu_state_device.argtypes = [_vp, C.POINTER(_vp), C.POINTER(_vp)]
    L.pepc_gpu_read_energies.argtypes = [_vp, _i64, _vp, C.POINTER(_i64)]
    L.pepc_gpu_adjoint_device.argtypes = [_vp] * 9
    L.pepc_gpu_autotune.argtypes = [_vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp, C.c_int32,
                                    C.POINTER(TuneEntry), _i64, C.POINTER(_i64)]
    L.pepc_gpu_get_config.argtypes = [_vp, C.POINTER(Config)]
    L.pepc_gpu_freeze.argtypes = [_vp, C.POINTER(_i64)]
    L.pepc_gpu_restore_device.argtypes = [_vp, _i64, C.POINTER(_vp), C.POINTER(_vp)]
    L.pepc_gpu_replay_device.argtypes = [_vp] * 9
    L.pepc_gpu_periodic_default.argtypes = [C.POINTER(Periodic)]
    L.pepc_gpu_periodic_default.restype = None
    L.pepc_gpu_set_periodic.argtypes = [_vp, C.POINTER(Periodic)]
    L.pepc_gpu_lattice_get.argtypes = [_vp, _vp, _vp, _vp, C.POINTER(C.c_int32)]
    L.pepc_gpu_constrain_periodic_device.argtypes = [_vp, _i64, _vp, _vp, _vp]
    L.pepc_gpu_traverse_sinks_device.argtypes = [_vp, _i64, _vp, _vp, _vp, C.c_int32] + [_vp] * 5
    L.pepc_gpu_traverse_sinks.argtypes = [_vp, _i64, _vp, _vp, _vp, C.c_int32] + [_vp] * 5
    L.pepc_gpu_read_sink_counts.argtypes = [_vp, _i64, _vp, _vp]
    L.pepc_gpu_traverse_sinks_noalloc.argtypes = [_vp, _i64, _vp, _vp, _vp, C.c_int32] + [_vp] * 5
    L.pepc_gpu_check_tree.argtypes = [_vp, C.POINTER(_i64)]
    L.pepc_gpu_reserve_nodes.argtypes = [_vp, _i64]
    L.pepc_gpu_read_box.argtypes = [_vp, _i64, _vp, _vp, _vp]
    for name in EXPORTS:
        f = getattr(L, name)
        if f.restype is not None and name not in ("pepc_gpu_config_default",
                                                  "pepc_gpu_periodic_default"):
            f.restype = C.c_int
    _lib = L
    return L


def default_config(**kw) -> Config:
    cfg = Config()
    load().pepc_gpu_config_default(C.byref(cfg))
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg
