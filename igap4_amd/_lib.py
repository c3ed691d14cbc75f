"""ctypes binding of libigap_gpu.so (the C ABI declared in include/igap_gpu.h and
include/igap_ws.h).

The library is built in-tree by igap4_amd/build.sh (or ``__graft_entry__.build()``).
There is no fallback: if the library or a GPU is missing, loading fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libigap_gpu.so")

IGAP_OK = 0
IGAP_ERR_ARG = 1
IGAP_ERR_HIP = 2
IGAP_ERR_NOMEM = 3
IGAP_ERR_STATE = 4
IGAP_PHYS_GRADELAST = 0
IGAP_PHYS_GRADELASTTIME = 2
IGAP_F64 = 0
IGAP_F32 = 1
IGAP_TANGENT_ELEMENT_MATRICES = 1
IGAP_TANGENT_DIRECT = 2
IGAP_GENERAL_GEOMETRY = 4
IGAP_PC_APPLY_FULL = 0
IGAP_PC_APPLY_SMOOTHER = 1

c_double_p = C.POINTER(C.c_double)
c_float_p = C.POINTER(C.c_float)
c_i64_p = C.POINTER(C.c_int64)
c_i32_p = C.POINTER(C.c_int32)


class MeshDesc(C.Structure):
    """igap_mesh_desc: struct __Mesh of one patch (include/meshimpl.h:6-29)."""
    _fields_ = [("porder", C.c_int), ("nelem", C.c_int * 3), ("periodic", C.c_int * 3),
                ("knot", c_double_p * 3), ("w", c_double_p), ("X", c_double_p)]


class PhysDesc(C.Structure):
    """igap_phys_desc: struct __Phys (include/physimpl.h:13-28)."""
    _fields_ = [("kind", C.c_int), ("par", C.c_double * 9), ("use_coef", C.c_int),
                ("coef", C.c_double * 6)]


class BCDesc(C.Structure):
    """igap_bc_desc: struct __BC (include/bcimpl.h:8-15)."""
    _fields_ = [("bc_type", C.c_int * 36), ("par_periodic", C.c_double * 9),
                ("par_dirichlet", C.c_double * 9), ("neumann", C.c_double * 36)]


class SolverOpts(C.Structure):
    """igap_solver_opts (IGAPSolvInit settings, src/interface.c:207-250)."""
    _fields_ = [("snes_atol", C.c_double), ("snes_rtol", C.c_double), ("snes_maxit", C.c_int),
                ("ksp_rtol", C.c_double), ("ksp_atol", C.c_double), ("ksp_maxit", C.c_int),
                ("ksp_restart", C.c_int), ("pc", C.c_int), ("tangent_prec", C.c_int),
                ("linesearch", C.c_int), ("pc_order", C.c_int),
                ("pc_lag", C.c_int), ("monitor", C.c_int), ("mg_levels", C.c_int),
                ("pc_fp32", C.c_int)]


class SolverStats(C.Structure):
    _fields_ = [("newton_its", C.c_int), ("ksp_its", C.c_int), ("ls_backtracks", C.c_int),
                ("n_residual", C.c_int), ("converged", C.c_int), ("fnorm0", C.c_double),
                ("fnorm", C.c_double), ("last_ksp_relres", C.c_double),
                ("t_residual_ms", C.c_double), ("t_tangent_ms", C.c_double), ("t_pc_ms", C.c_double),
                ("t_ksp_ms", C.c_double), ("t_total_ms", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


IGAP_PC_ILU0 = 0
IGAP_PC_JACOBI = 1
IGAP_PC_NONE = 2
IGAP_PC_MG = 3

# every symbol include/igap_gpu.h declares, with its ctypes signature
_SIGS = {
    "igap_phys_gradelast": (None, [C.POINTER(PhysDesc)]),
    "igap_phys_gradelasttime": (None, [C.POINTER(PhysDesc)]),
    "igap_bc_default": (None, [C.POINTER(BCDesc)]),
    "igap_mesh_igabeam": (C.c_int, [c_double_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int,
                                    c_double_p, c_double_p, c_double_p, c_double_p, c_double_p]),
    "igap_gpu_create": (C.c_int, [C.POINTER(MeshDesc), C.POINTER(PhysDesc), C.POINTER(BCDesc),
                                  C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "igap_gpu_destroy": (C.c_int, [C.c_void_p]),
    "igap_gpu_last_error": (C.c_char_p, []),
    "igap_gpu_set_bc": (C.c_int, [C.c_void_p, C.POINTER(BCDesc)]),
    "igap_gpu_sizes": (C.c_int, [C.c_void_p, c_i64_p, c_i64_p, c_i64_p, C.POINTER(C.c_int)]),
    "igap_gpu_uhist": (C.c_void_p, [C.c_void_p]),
    "igap_gpu_intdensity": (C.c_int, [C.c_void_p, c_double_p, C.POINTER(C.c_int)]),
    "igap_gpu_save_solution": (C.c_int, [C.c_void_p, C.c_char_p]),
    "igap_gpu_load_solution": (C.c_int, [C.c_void_p, C.c_char_p]),
    "igap_gpu_csr_rowptr": (C.c_void_p, [C.c_void_p]),
    "igap_gpu_csr_col": (C.c_void_p, [C.c_void_p]),
    "igap_gpu_csr_val64": (C.c_void_p, [C.c_void_p]),
    "igap_gpu_csr_val32": (C.c_void_p, [C.c_void_p]),
    "igap_gpu_dirichlet_mask": (C.c_void_p, [C.c_void_p]),
    "igap_gpu_dirichlet_values": (C.c_void_p, [C.c_void_p]),
    "igap_gpu_residual": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "igap_gpu_rvec": (C.c_void_p, [C.c_void_p]),
    "igap_gpu_tangent": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    "igap_gpu_synchronize": (C.c_int, [C.c_void_p]),
    "igap_gpu_malloc": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_int64]),
    "igap_gpu_free": (C.c_int, [C.c_void_p, C.c_void_p]),
    "igap_gpu_memcpy_htod": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]),
    "igap_gpu_memcpy_dtoh": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]),
    "igap_gpu_memcpy_dtod": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]),
    "igap_gpu_timing_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "igap_gpu_timing_get": (C.c_int, [C.c_void_p, C.c_int, c_double_p, c_i64_p]),
    "igap_gpu_timing_reset": (C.c_int, [C.c_void_p]),
    "igap_gpu_selftest_mfma": (C.c_int, [c_double_p, c_float_p]),
    "igap_gpu_debug_fail_alloc": (C.c_int, [C.c_int]),
    "igap_gpu_residual_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "igap_gpu_phys_eval": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "igap_gpu_selftest_ilu": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), c_double_p]),
    "igap_gpu_solver_default_opts": (C.c_int, [C.POINTER(SolverOpts)]),
    "igap_gpu_snes_solve": (C.c_int, [C.c_void_p, C.POINTER(SolverOpts), C.POINTER(SolverStats)]),
    "igap_gpu_pc_apply": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "igap_gpu_tangent_apply": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    # include/igap_ws.h
    "igap_ws_create": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.c_int, C.c_int, C.POINTER(C.c_int), C.c_uint64,
                                 C.POINTER(C.c_void_p)]),
    "igap_ws_destroy": (C.c_int, [C.c_void_p]),
    "igap_ws_nparams": (C.c_int, [C.c_void_p, c_i64_p]),
    "igap_ws_get_params": (C.c_int, [C.c_void_p, c_float_p]),
    "igap_ws_set_params": (C.c_int, [C.c_void_p, c_float_p]),
    "igap_ws_forward": (C.c_int, [C.c_void_p, c_float_p, c_float_p]),
    "igap_ws_gradients": (C.c_int, [C.c_void_p, c_float_p, c_float_p, c_double_p, c_float_p]),
    "igap_ws_train_step": (C.c_int, [C.c_void_p, c_float_p, c_float_p, C.c_float, c_double_p]),
    "igap_ws_selftest_mfma_bf16": (C.c_int, [c_float_p]),
}

_LIB = None


class IgapError(RuntimeError):
    pass


def lib():
    """Load libigap_gpu.so; raises if it is not built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise IgapError(f"{LIB_PATH} is not built: run igap4_amd/build.sh "
                            "(or __graft_entry__.build()); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def declared_symbols():
    return list(_SIGS)


def check(rc):
    if rc != IGAP_OK:
        msg = lib().igap_gpu_last_error()
        raise IgapError(f"igap error {rc}: {msg.decode() if msg else ''}")
