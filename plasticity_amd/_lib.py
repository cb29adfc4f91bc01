"""ctypes binding of libpfem.so (the C-ABI declared in include/pfem.h).

The library is built in-tree by ``plasticity_amd/build.py`` (``__graft_entry__.build()``).
Loading fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# PFEM_LIB selects another build of the same library (e.g. a check build)
LIBPATH = os.environ.get("PFEM_LIB") or os.path.join(HERE, "libpfem.so")

PFEM_OK = 0
PFEM_ERR_INVERTED = 1
PFEM_ERR_NOT_CONVERGED = 2
PFEM_ERR_NON_SPD = 3
PFEM_ERR_UNSUPPORTED = 4
PFEM_ERR_HIP = 5
PFEM_ERR_BAD_ARG = 6
PFEM_ERR_NO_DEVICE = 7
BATCH_NOALLOC_MAX = 2048

HEX8_FINITE = 0
HEX8_FINITE_FBAR = 1
HEX8_SMALL = 2
PREC_FP64 = 0
PREC_FP32_IR = 1
SMOOTH_JACOBI = 0
SMOOTH_CHEBYSHEV = 1
SMOOTH_CHEBYSHEV4 = 2


class Material(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("E", "nu", "sy0", "Hiso", "Hkin", "ssat", "delta")]


class MeshDesc(C.Structure):
    _fields_ = [("nnodes", C.c_int64), ("nelem", C.c_int64), ("nodes", C.c_void_p),
                ("conn", C.c_void_p), ("material_id", C.c_void_p), ("nmaterials", C.c_int32),
                ("materials", C.c_void_p), ("element_kind", C.c_int32)]


AMG_PER_LOAD_STEP = 0
AMG_EVERY_ITERATION = 1
AMG_FIRST_TWO = 2
AMG_FINE_SMOOTHER = 3
AMG_KEEP_P_IN_STEP = 4
AMG_KEEP_P = 5
AMG_KEEP_P_UNTIL_FALLBACK = 6
AMG_REUSE_COARSE = 7
BC_RAMPED = 0
BC_FIXED = 1


class Options(C.Structure):
    _fields_ = [("precision", C.c_int32), ("smoother", C.c_int32), ("smoother_degree", C.c_int32),
                ("cg_itmax", C.c_int32), ("eta_min", C.c_double), ("eta_max", C.c_double),
                ("ew_gamma", C.c_double), ("ew_alpha", C.c_double), ("coarse_max", C.c_int32),
                ("max_levels", C.c_int32), ("verbose", C.c_int32), ("amg_rebuild", C.c_int32),
                ("amg_wcycle", C.c_int32), ("krylov_fallback", C.c_int32), ("gcr_directions", C.c_int32),
                ("fp16_fine", C.c_int32), ("line_search", C.c_int32), ("rho_iters_cold", C.c_int32),
                ("rho_iters_warm", C.c_int32), ("host_build", C.c_int32), ("vcycle_eager", C.c_int32),
                ("prolong_rowwise", C.c_int32), ("fp32_krylov", C.c_int32), ("newton_diverge", C.c_double),
                ("assembly_chunk", C.c_int32), ("smoother_unfused", C.c_int32),
                ("coarse_degree", C.c_int32), ("krylov_sync", C.c_int32),
                ("split_operator", C.c_int32), ("reserved", C.c_int32 * 3)]


class Fail(C.Structure):
    _fields_ = [("e", C.c_int64), ("g", C.c_int32), ("J", C.c_double), ("hip", C.c_int32)]


class NewtonResult(C.Structure):
    _fields_ = [("converged", C.c_int32), ("iters", C.c_int32),
                ("krylov_iters", C.c_int32 * 64), ("solver_used", C.c_int32 * 64),
                ("residuals", C.c_double * 64), ("t_assemble", C.c_double), ("t_bc", C.c_double),
                ("t_setup", C.c_double), ("t_solve", C.c_double), ("t_total", C.c_double),
                ("eta", C.c_double * 64), ("solves", C.c_int32), ("discarded", C.c_int32),
                ("stopped", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("nnodes", C.c_int64), ("nelem", C.c_int64), ("ndof", C.c_int64), ("nnzb", C.c_int64),
                ("ncolors", C.c_int64), ("amg_levels", C.c_int32), ("amg_ndof", C.c_int64 * 12),
                ("amg_nnzb", C.c_int64 * 12), ("amg_operator_complexity", C.c_double),
                ("t_assemble", C.c_double), ("t_setup", C.c_double), ("t_solve", C.c_double),
                ("t_gp_kernel", C.c_double), ("t_scatter_kernel", C.c_double), ("t_spmv", C.c_double),
                ("t_pcg_spmv", C.c_double), ("n_pcg_spmv", C.c_int64),
                ("t_setup_rho", C.c_double), ("t_setup_galerkin", C.c_double), ("t_setup_coarse", C.c_double),
                ("n_assemble", C.c_int64), ("n_setup", C.c_int64), ("n_solve", C.c_int64),
                ("n_krylov", C.c_int64), ("device_bytes", C.c_int64), ("t_create_pattern", C.c_double),
                ("built_on_device", C.c_int32), ("t_krylov_spmv32", C.c_double), ("n_krylov_spmv32", C.c_int64)]


class NetDesc(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("channels", C.c_int32 * 4),
                ("capacity", C.c_int64), ("seed", C.c_uint64)]


# symbol -> (restype, argtypes); every declaration of include/pfem.h
_vp = C.c_void_p
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_dp = C.POINTER(C.c_double)
SIGNATURES = {
    "pfem_default_options": (None, [C.POINTER(Options)]),
    "pfem_create": (C.c_int, [C.POINTER(MeshDesc), C.POINTER(Options), C.POINTER(_vp)]),
    "pfem_box_mesh": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_int64, C.c_int64, C.c_int64, C.c_double,
                                C.c_uint64, _vp, _vp]),
    "pfem_destroy": (None, [_vp]),
    "pfem_set_dirichlet": (C.c_int, [_vp, C.c_int64, _vp, _vp, _vp]),
    "pfem_set_neumann": (C.c_int, [_vp, C.c_int64, _vp, _vp]),
    "pfem_reset": (C.c_int, [_vp]),
    "pfem_set_displacement": (C.c_int, [_vp, _vp]),
    "pfem_get_displacement": (C.c_int, [_vp, _vp]),
    "pfem_set_state": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "pfem_get_state": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "pfem_get_trial_stress": (C.c_int, [_vp, _vp]),
    "pfem_gauss_stress": (C.c_int, [_vp, _vp]),
    "pfem_rollback": (C.c_int, [_vp]),
    "pfem_assemble": (C.c_int, [_vp, _vp, C.c_int, _vp, C.POINTER(Fail)]),
    "pfem_subtract_external": (C.c_int, [_vp, C.c_double]),
    "pfem_impose_dirichlet": (C.c_int, [_vp, C.c_double]),
    "pfem_impose_dirichlet_set": (C.c_int, [_vp, C.c_int, C.c_double, _vp, _vp]),
    "pfem_residual_norm": (C.c_int, [_vp, _dp]),
    "pfem_linear_solve": (C.c_int, [_vp, _vp, C.c_double, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), _dp,
                                    _vp]),
    "pfem_update_displacement": (C.c_int, [_vp]),
    "pfem_get_K_bsr": (C.c_int, [_vp, _i64p, _vp, _vp, _vp]),
    "pfem_get_K_csc": (C.c_int, [_vp, _i64p, _vp, _vp, _vp]),
    "pfem_get_R": (C.c_int, [_vp, _vp]),
    "pfem_get_dU": (C.c_int, [_vp, _vp]),
    "pfem_get_colors": (C.c_int, [_vp, _vp, _i32p]),
    "pfem_newton": (C.c_int, [_vp, C.c_double, C.c_double, C.c_int, C.POINTER(NewtonResult), C.POINTER(Fail)]),
    "pfem_commit": (C.c_int, [_vp, C.POINTER(Fail)]),
    "pfem_element_batch": (C.c_int, [C.c_int32, C.POINTER(Material), C.c_int64, _vp, _vp, _vp, _vp, _vp, _vp,
                                     C.c_int, _vp, _vp, _vp, C.POINTER(Fail)]),
    "pfem_element_batch_noalloc": (C.c_int, [C.c_int32, C.POINTER(Material), C.c_int64, _vp, _vp, _vp, _vp, _vp,
                                             _vp, C.c_int, _vp, _vp, _vp, C.POINTER(Fail)]),
    "pfem_get_stats": (C.c_int, [_vp, C.POINTER(Stats)]),
    "pfem_bench_spmv": (C.c_int, [_vp, C.c_int, C.c_int, _dp]),
    "pfem_apply_K": (C.c_int, [_vp, C.c_int, _dp, _dp]),
    "pfem_bench_vcycle": (C.c_int, [_vp, C.c_int, C.c_int, _dp, _dp, _dp, _dp]),
    "pfem_dense_inverse": (C.c_int, [C.c_int64, _vp, _vp]),
    "pfem_device_peaks": (C.c_int, [C.c_int64, C.c_int, _dp, _dp, _dp]),
    "pfem_status_string": (C.c_char_p, [C.c_int]),
    "pfem_net_create": (C.c_int, [C.POINTER(NetDesc), C.POINTER(_vp)]),
    "pfem_net_destroy": (None, [_vp]),
    "pfem_net_set_conv_kernel": (C.c_int, [_vp, C.c_int]),
    "pfem_net_set_loss": (C.c_int, [_vp, C.c_int]),
    "pfem_net_num_params": (C.c_int64, [_vp]),
    "pfem_net_get_params": (C.c_int, [_vp, _vp]),
    "pfem_net_set_params": (C.c_int, [_vp, _vp]),
    "pfem_net_forward": (C.c_int, [_vp, _vp, _vp]),
    "pfem_net_loss_grad": (C.c_int, [_vp, _vp, _vp, _dp, _vp]),
    "pfem_net_add_pair": (C.c_int, [_vp, _vp, _vp]),
    "pfem_net_num_pairs": (C.c_int, [_vp, _i64p]),
    "pfem_net_train": (C.c_int, [_vp, C.c_int, C.c_int, C.c_double, C.c_uint64, _dp]),
    "pfem_attach_net": (C.c_int, [_vp, _vp, C.c_int, C.c_int]),
}

_lib = None


def lib():
    """Load libpfem.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise ImportError(f"{LIBPATH} is missing: build it with plasticity_amd/build.py "
                              "(or __graft_entry__.build()); there is no CPU fallback")
        L = C.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class PfemError(RuntimeError):
    def __init__(self, status, msg=""):
        s = lib().pfem_status_string(status).decode()
        super().__init__(f"{msg}: {s} (status {status})" if msg else f"{s} (status {status})")
        self.status = status


def check(rc, what=""):
    if rc != PFEM_OK:
        raise PfemError(rc, what)


def ptr(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(C.c_void_p)
