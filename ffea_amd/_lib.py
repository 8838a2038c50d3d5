"""ctypes binding of libffea_hip.so (include/ffea_hip.h).

The library is the only compute path: if it is missing or cannot be loaded this
module raises, there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libffea_hip.so")

FFH_OK = 0
FFH_E_ARG = -1
FFH_E_HIP = -2
FFH_E_INVERSION = -3
FFH_E_CG = -4
FFH_E_NOMEM = -5
FFH_PRECISION_FP64 = 0
FFH_PRECISION_MIXED = 1
FFH_NUM_KERNELS = 15
KERNEL_NAMES = ("element", "assemble", "rhs", "cg_spmv", "cg_update", "integrate", "mpnn_edges",
                "mpnn_embed", "mpnn_message", "mpnn_dense", "mpnn_out", "cg_totals", "ssint_forces", "ssint_grid",
                "box_wrap")

# Every symbol include/ffea_hip.h declares.
EXPORTS = (
    "ffh_last_error", "ffh_version", "ffh_blob_create", "ffh_blob_destroy",
    "ffh_blob_update_internal_forces", "ffh_blob_update_positions", "ffh_blob_step",
    "ffh_blob_solve", "ffh_blob_apply_matrix", "ffh_blob_set_positions",
    "ffh_blob_get_positions", "ffh_blob_get_velocities", "ffh_blob_get_rhs",
    "ffh_blob_get_rng_state", "ffh_blob_set_rng_state", "ffh_blob_last_cg", "ffh_blob_sizes",
    "ffh_blob_stream", "ffh_blob_synchronize", "ffh_blob_set_profiling",
    "ffh_blob_kernel_times", "ffh_blob_strain_energy", "ffh_batch_create",
    "ffh_batch_num_blobs", "ffh_batch_last_cg", "ffh_mpnn_num_params", "ffh_mpnn_init",
    "ffh_mpnn_set_weights", "ffh_mpnn_get_weights", "ffh_blob_set_preconditioner",
    "ffh_mpnn_apply", "ffh_mpnn_harvest", "ffh_mpnn_harvest_count", "ffh_mpnn_train",
    "ffh_mpnn_loss_grad", "ffh_blob_steps_done", "ffh_blob_set_graphs", "ffh_blob_tuning",
    "ffh_blob_set_tuning", "ffh_blob_set_external_forces", "ffh_blob_set_solver",
    "ffh_blob_set_velocities", "ffh_blob_kinetic_energy", "ffh_blob_set_springs",
    "ffh_blob_spring_energy", "ffh_blob_set_constant_forces", "ffh_blob_set_surface_forces",
    "ffh_blob_set_rotational_forces", "ffh_blob_set_box", "ffh_blob_set_ssint", "ffh_blob_set_steric", "ffh_blob_ssint_energy",
    "ffh_blob_ssint_energy_noalloc", "ffh_blob_create_meso", "ffh_blob_update_matrix",
    "ffh_blob_set_stress_output", "ffh_blob_get_stress_magnitudes",
)
FFH_PRECOND_JACOBI = 0
FFH_PRECOND_LEARNED = 1
# solver codes of include/FFEA_return_codes.h:30-33, by .ffea "solver" value
FFH_SOLVERS = {"CG": 1, "masslumped": 2, "CG_nomass": 3}
# surface-surface interaction types on the device, by .ffea ssint_type value
FFH_SSINT_KINDS = {"lennard-jones": 1, "ljsteric": 2, "gensoft": 3, "steric": 4}

_dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")

_lib = None


class FFHError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ffea_hip error {code}: {msg}")
        self.code = code


class InversionError(FFHError):
    pass


class CGError(FFHError):
    pass


def load(path: str | None = None) -> C.CDLL:
    """Load libffea_hip.so (raises if absent: build it with `python -m ffea_amd.build`)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(f"{p} not built; run `python -m ffea_amd.build`")
    L = C.CDLL(p)
    L.ffh_last_error.restype = C.c_char_p
    L.ffh_version.restype = C.c_int
    L.ffh_blob_create.restype = C.c_void_p
    L.ffh_blob_create.argtypes = [C.c_int, C.c_int, _dp, _ip, _dp, _dp, C.c_int, _ip, C.c_double,
                                  C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_double,
                                  _u64p, C.c_int]
    L.ffh_blob_create_meso.restype = C.c_void_p
    L.ffh_blob_create_meso.argtypes = L.ffh_blob_create.argtypes
    L.ffh_batch_create.restype = C.c_void_p
    L.ffh_batch_create.argtypes = [C.c_int, _ip, _ip, _dp, _ip, _dp, _dp, _ip, _ip, C.c_double,
                                   C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_double,
                                   _u64p, C.c_int]
    L.ffh_batch_num_blobs.argtypes = [C.c_void_p]
    L.ffh_batch_last_cg.argtypes = [C.c_void_p, _ip, _dp, _ip]
    _fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    L.ffh_mpnn_num_params.argtypes = []
    L.ffh_mpnn_init.argtypes = [C.c_void_p, C.c_ulonglong]
    L.ffh_mpnn_set_weights.argtypes = [C.c_void_p, _fp]
    L.ffh_mpnn_get_weights.argtypes = [C.c_void_p, _fp]
    L.ffh_blob_set_preconditioner.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.ffh_mpnn_apply.argtypes = [C.c_void_p, _dp, _dp]
    L.ffh_mpnn_harvest.argtypes = [C.c_void_p, C.c_int]
    L.ffh_mpnn_harvest_count.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_longlong)]
    L.ffh_mpnn_train.argtypes = [C.c_void_p, C.c_int, C.c_float, C.c_ulonglong, _fp]
    L.ffh_mpnn_loss_grad.argtypes = [C.c_void_p, _dp, _dp, C.POINTER(C.c_double), _fp]
    L.ffh_blob_steps_done.argtypes = [C.c_void_p, C.POINTER(C.c_longlong)]
    L.ffh_blob_set_graphs.argtypes = [C.c_void_p, C.c_int]
    L.ffh_blob_set_external_forces.argtypes = [C.c_void_p, C.c_void_p]
    L.ffh_blob_tuning.argtypes = [C.c_void_p] + [C.POINTER(C.c_int)] * 5
    L.ffh_blob_set_tuning.argtypes = [C.c_void_p] + [C.c_int] * 5
    L.ffh_blob_destroy.argtypes = [C.c_void_p]
    L.ffh_blob_destroy.restype = None
    L.ffh_blob_update_internal_forces.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
    for n in ("ffh_blob_update_positions", "ffh_blob_synchronize"):
        getattr(L, n).argtypes = [C.c_void_p]
    L.ffh_blob_step.argtypes = [C.c_void_p, C.c_int]
    L.ffh_blob_solve.argtypes = [C.c_void_p, _dp]
    L.ffh_blob_apply_matrix.argtypes = [C.c_void_p, _dp, _dp]
    for n in ("ffh_blob_set_positions", "ffh_blob_get_positions", "ffh_blob_get_velocities",
              "ffh_blob_get_rhs"):
        getattr(L, n).argtypes = [C.c_void_p, _dp]
    L.ffh_blob_get_rng_state.argtypes = [C.c_void_p, _u64p]
    L.ffh_blob_set_rng_state.argtypes = [C.c_void_p, _u64p]
    L.ffh_blob_last_cg.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.ffh_blob_sizes.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                 C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    L.ffh_blob_stream.restype = C.c_void_p
    L.ffh_blob_stream.argtypes = [C.c_void_p]
    L.ffh_blob_set_profiling.argtypes = [C.c_void_p, C.c_int]
    L.ffh_blob_kernel_times.argtypes = [C.c_void_p, _dp, _i64p]
    L.ffh_blob_strain_energy.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    L.ffh_blob_set_solver.argtypes = [C.c_void_p, C.c_int]
    L.ffh_blob_set_velocities.argtypes = [C.c_void_p, _dp]
    L.ffh_blob_kinetic_energy.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    L.ffh_blob_set_springs.argtypes = [C.c_void_p, C.c_int, _ip, _ip, _dp, _dp]
    L.ffh_blob_spring_energy.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    L.ffh_blob_set_constant_forces.argtypes = [C.c_void_p, C.c_int, _ip, _dp]
    L.ffh_blob_set_rotational_forces.argtypes = [C.c_void_p, C.c_int, _ip, _dp, _dp, _ip]
    L.ffh_blob_set_box.argtypes = [C.c_void_p, _dp, C.c_int]
    L.ffh_blob_set_ssint.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip, _ip, _ip, C.c_int, _dp, _dp, _dp, _ip,
                                     C.c_double, C.c_int, C.c_int]
    L.ffh_blob_set_steric.argtypes = [C.c_void_p, C.c_double, C.c_double]
    L.ffh_blob_ssint_energy.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    L.ffh_blob_ssint_energy_noalloc.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    L.ffh_blob_update_matrix.argtypes = [C.c_void_p]
    L.ffh_blob_set_stress_output.argtypes = [C.c_void_p, C.c_int]
    L.ffh_blob_get_stress_magnitudes.argtypes = [C.c_void_p, _dp]
    L.ffh_blob_set_surface_forces.argtypes = [C.c_void_p, C.c_int, _ip, C.c_int, _ip, _ip, _dp]
    for n in EXPORTS:
        if n not in ("ffh_last_error", "ffh_version", "ffh_blob_create", "ffh_blob_destroy",
                     "ffh_blob_stream", "ffh_batch_create", "ffh_blob_create_meso"):
            getattr(L, n).restype = C.c_int
    if path is None:
        _lib = L
    return L


def last_error() -> str:
    return load().ffh_last_error().decode()


def check(rc: int) -> int:
    if rc == FFH_OK:
        return rc
    msg = last_error()
    if rc == FFH_E_INVERSION:
        raise InversionError(rc, msg)
    if rc == FFH_E_CG:
        raise CGError(rc, msg)
    raise FFHError(rc, msg)
