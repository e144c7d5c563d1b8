"""ctypes binding of libimps.so (the C ABI declared in include/imps.h).

The library is built in-tree by ``infinitempsalgorithms_amd.build``.  There is no CPU
fallback: every entry point of this package calls the HIP library and raises when it
is missing or fails.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libimps.so")

IMPS_OK = 0
IMPS_E_ARG, IMPS_E_STATE, IMPS_E_HIP, IMPS_E_SOLVER, IMPS_E_UNSUPPORTED, IMPS_E_IO = -1, -2, -3, -4, -5, -6
IMPS_PREC_FP64, IMPS_PREC_FP32, IMPS_PREC_BF16X3 = 0, 1, 2
IMPS_SIDE_LEFT, IMPS_SIDE_RIGHT = 0, 1

c_int, c_double, c_char, c_void_p, c_size_t = (ctypes.c_int, ctypes.c_double, ctypes.c_char,
                                               ctypes.c_void_p, ctypes.c_size_t)
P_double = ctypes.POINTER(ctypes.c_double)
PP_double = ctypes.POINTER(P_double)
P_int = ctypes.POINTER(ctypes.c_int)


class KrylovTiming(ctypes.Structure):
    _fields_ = [("ms_multidot", c_double), ("ms_axpy_dots", c_double), ("ms_axpy_norm", c_double),
                ("bytes_multidot", c_double), ("bytes_axpy", c_double)]


class KrylovParams(ctypes.Structure):
    _fields_ = [("tol", c_double), ("krylovdim", c_int), ("maxiter", c_int), ("eager", c_int),
                ("prec", c_int)]


class KrylovInfo(ctypes.Structure):
    _fields_ = [("converged", c_int), ("numiter", c_int), ("numops", c_int), ("normres", c_double)]


class ResponseInfo(ctypes.Structure):
    _fields_ = [("iterations", c_int), ("converged", c_int), ("relres", c_double), ("estimate", c_double),
                ("hessian_applies", ctypes.c_long), ("level_solves", ctypes.c_long), ("level_ops", ctypes.c_long),
                ("unconverged_solves", ctypes.c_long)]


class BatchStats(ctypes.Structure):
    _fields_ = [("rounds", ctypes.c_long), ("launches", ctypes.c_long), ("seconds", c_double), ("groups", c_int),
                ("t_phase", c_double * 6)]


FINALIZE_CB = ctypes.CFUNCTYPE(c_int, c_int, c_void_p, c_void_p)


class VumpsParams(ctypes.Structure):
    _fields_ = [("tol", c_double), ("maxiter", c_int), ("verbosity", c_int), ("tol_min", c_double),
                ("tol_max", c_double), ("eigs_tolfactor", c_double), ("gauge_tolfactor", c_double),
                ("envs_tolfactor", c_double), ("krylovdim", c_int), ("eig_maxiter", c_int),
                ("gauge_maxiter", c_int), ("envs_maxiter", c_int), ("prec", c_int),
                ("refine_below", c_double), ("eig_floor", c_double),
                ("eig_dynamic", c_int), ("eig_eager", c_int), ("eig_tol", c_double),
                ("gauge_dynamic", c_int), ("gauge_tol", c_double), ("gauge_tol_min", c_double),
                ("gauge_tol_max", c_double), ("envs_dynamic", c_int), ("envs_tol", c_double),
                ("envs_tol_min", c_double), ("envs_tol_max", c_double), ("envs_krylovdim", c_int),
                ("finalize", FINALIZE_CB), ("finalize_user", c_void_p), ("refine_eig", c_int),
                ("gauge_eig_miniter", c_int)]


class VumpsInfo(ctypes.Structure):
    _fields_ = [("niter", c_int), ("converged", c_int), ("itererr", c_double), ("energy", c_double),
                ("seconds", c_double), ("t_galerkin_reached", c_double), ("hac_applies", c_int),
                ("hc_applies", c_int), ("prec_switch_iter", c_int), ("t_local", c_double),
                ("t_gauge", c_double), ("t_envs", c_double), ("t_galerkin", c_double), ("t_energy", c_double),
                ("gauge_sweeps", c_int), ("gmres_ops", c_int), ("qr_fallbacks", c_int),
                ("refine_fallbacks", c_int)]


class HacTiming(ctypes.Structure):
    _fields_ = [("ms_apply", c_double), ("ms_gemm1", c_double), ("ms_wmix", c_double),
                ("ms_gemm2", c_double), ("flop_gemm1", c_double), ("flop_gemm2", c_double),
                ("ms_prepare", c_double)]


class BatchInstance(ctypes.Structure):
    _fields_ = [("nsites", c_int), ("w", c_int), ("d", c_int), ("D", c_int), ("W", PP_double), ("A", PP_double),
                ("status", c_int), ("niter", c_int), ("converged", c_int), ("itererr", c_double),
                ("energy", c_double), ("seconds", c_double), ("hac_applies", c_int)]


ITER_CB = ctypes.CFUNCTYPE(c_int, c_int, c_double, c_double, c_void_p)

# (name, restype, argtypes) of every symbol declared in include/imps.h
SIGNATURES = [
    ("imps_create", c_int, [c_int, ctypes.POINTER(c_void_p)]),
    ("imps_destroy", c_int, [c_void_p]),
    ("imps_last_error", ctypes.c_char_p, [c_void_p]),
    ("imps_set_mpo", c_int, [c_void_p, c_int, c_int, c_int, PP_double]),
    ("imps_set_mpo_chain", c_int, [c_void_p, c_int, P_int, P_int, PP_double]),
    ("imps_set_state", c_int, [c_void_p, c_int, c_int, PP_double, PP_double, PP_double, PP_double]),
    ("imps_get_state", c_int, [c_void_p, PP_double, PP_double, PP_double, PP_double]),
    ("imps_state_dims", c_int, [c_void_p, P_int, P_int, P_int]),
    ("imps_bench_krylov", c_int, [c_void_p, ctypes.c_long, c_int, c_int, ctypes.POINTER(KrylovTiming)]),
    ("imps_write_state_file", c_int, [ctypes.c_char_p, c_int, c_int, c_int, PP_double, PP_double, PP_double,
                                      PP_double]),
    ("imps_state_file_info", c_int, [ctypes.c_char_p, P_int, P_int, P_int]),
    ("imps_read_state_file", c_int, [ctypes.c_char_p, c_int, c_int, c_int, PP_double, PP_double, PP_double,
                                     PP_double]),
    ("imps_write_mpo_file", c_int, [ctypes.c_char_p, c_int, c_int, c_int, PP_double]),
    ("imps_mpo_file_info", c_int, [ctypes.c_char_p, P_int, P_int, P_int]),
    ("imps_read_mpo_file", c_int, [ctypes.c_char_p, c_int, c_int, c_int, PP_double]),
    ("imps_checkpoint_state", c_int, [c_void_p, ctypes.c_char_p]),
    ("imps_resume_state", c_int, [c_void_p, ctypes.c_char_p]),
    ("imps_set_state_raw", c_int, [c_void_p, c_int, c_int, PP_double, c_double, c_int]),
    ("imps_hac_apply", c_int, [c_void_p, c_int, P_double, P_double, c_int]),
    ("imps_hac_apply_dev", c_int, [c_void_p, c_int, c_void_p, c_void_p, c_int]),
    ("imps_hac_apply_noalloc", c_int, [c_void_p, c_int, c_void_p, c_void_p, c_int]),
    ("imps_hc_apply", c_int, [c_void_p, c_int, P_double, P_double, c_int]),
    ("imps_hc_apply_dev", c_int, [c_void_p, c_int, c_void_p, c_void_p, c_int]),
    ("imps_transfer_apply", c_int, [c_void_p, c_int, P_double, P_double]),
    ("imps_regularize", c_int, [c_void_p, c_int, P_double, P_double, P_double]),
    ("imps_envs_recalculate", c_int, [c_void_p, c_double, c_int, c_int]),
    ("imps_envs_reset", c_int, [c_void_p]),
    ("imps_get_env", c_int, [c_void_p, c_int, c_int, P_double]),
    ("imps_set_env", c_int, [c_void_p, c_int, c_int, P_double]),
    ("imps_eigsolve_ac", c_int, [c_void_p, c_int, ctypes.POINTER(KrylovParams), P_double, P_double,
                                 ctypes.POINTER(KrylovInfo)]),
    ("imps_eigsolve_c", c_int, [c_void_p, c_int, ctypes.POINTER(KrylovParams), P_double, P_double,
                                ctypes.POINTER(KrylovInfo)]),
    ("imps_regauge", c_int, [c_void_p, c_int, P_double]),
    ("imps_gaugefix_right", c_int, [c_void_p, c_double, c_int, ctypes.POINTER(c_int)]),
    ("imps_galerkin", c_int, [c_void_p, P_double]),
    ("imps_energy", c_int, [c_void_p, P_double]),
    ("imps_energy_gradient", c_int, [c_void_p, PP_double]),
    ("imps_fixedpoint_gradient", c_int, [c_void_p, c_int, c_int, P_double, c_double, c_int, c_double, P_double,
                                         P_double, ctypes.POINTER(ResponseInfo)]),
    ("imps_tangent_hessian_apply", c_int, [c_void_p, c_double, c_double, P_double, P_double]),
    ("imps_local_expectation", c_int, [c_void_p, c_int, P_double, P_double]),
    ("imps_local_expectation_complex", c_int, [c_void_p, c_int, P_double, P_double]),
    ("imps_vumps_default_params", c_int, [ctypes.POINTER(VumpsParams)]),
    ("imps_vumps_run", c_int, [c_void_p, ctypes.POINTER(VumpsParams), ctypes.POINTER(VumpsInfo), P_double,
                               c_int, ITER_CB, c_void_p]),
    ("imps_idmrg_run", c_int, [c_void_p, ctypes.POINTER(VumpsParams), ctypes.POINTER(VumpsInfo), P_double,
                               c_int, ITER_CB, c_void_p]),
    ("imps_tdvp_timestep", c_int, [c_void_p, c_double, c_double, c_int, c_int, c_double, c_int]),
    ("imps_tdvp_timestep_complex", c_int, [c_void_p, c_double, c_double, c_double, c_int, c_int, c_double, c_int]),
    ("imps_set_state_complex", c_int, [c_void_p, c_int, c_int, PP_double, PP_double, PP_double, PP_double]),
    ("imps_set_state_raw_complex", c_int, [c_void_p, c_int, c_int, PP_double, c_double, c_int]),
    ("imps_get_state_complex", c_int, [c_void_p, PP_double, PP_double, PP_double, PP_double]),
    ("imps_state_is_complex", c_int, [c_void_p, P_int]),
    ("imps_set_state_nonuniform", c_int, [c_void_p, c_int, c_int, P_int, PP_double, PP_double, PP_double, PP_double]),
    ("imps_set_state_raw_nonuniform", c_int, [c_void_p, c_int, c_int, P_int, PP_double, c_double, c_int]),
    ("imps_set_state_complex_nonuniform", c_int,
     [c_void_p, c_int, c_int, P_int, PP_double, PP_double, PP_double, PP_double]),
    ("imps_set_state_raw_complex_nonuniform", c_int, [c_void_p, c_int, c_int, P_int, PP_double, c_double, c_int]),
    ("imps_bond_dims", c_int, [c_void_p, P_int]),
    ("imps_correlator", c_int, [c_void_p, P_double, P_double, c_int, c_int, c_void_p, P_double]),
    ("imps_correlator_complex", c_int, [c_void_p, P_double, P_double, c_int, c_int, c_void_p, P_double]),
    ("imps_dot_state", c_int, [c_void_p, c_int, PP_double, P_double]),
    ("imps_dot_state_complex", c_int, [c_void_p, c_int, PP_double, P_double]),
    ("imps_dominant_env", c_int, [c_void_p, c_int, c_int, c_int, c_int, PP_double, c_double, c_int, c_int, P_double,
                                  P_double]),
    ("imps_dense_expectation", c_int, [c_void_p, c_int, c_int, c_int, PP_double, c_double, c_int, c_int, P_double,
                                       PP_double, PP_double]),
    ("imps_vumps_batch_device", c_int, [c_int, c_int, ctypes.POINTER(BatchInstance), ctypes.POINTER(VumpsParams),
                                        c_double, ctypes.POINTER(BatchStats)]),
    ("imps_vumps_batch", c_int, [c_int, c_int, ctypes.POINTER(BatchInstance), ctypes.POINTER(VumpsParams), c_int,
                                 c_double]),
    ("imps_bench_hac", c_int, [c_void_p, c_int, c_int, c_int, ctypes.POINTER(HacTiming)]),
    ("imps_set_gemm_autotune", c_int, [c_int]),
    ("imps_set_graph_capture", c_int, [c_int]),
    ("imps_gemm_dev", c_int, [c_void_p, c_char, c_char, c_int, c_int, c_int, c_double, c_void_p, c_int,
                              c_void_p, c_int, c_double, c_void_p, c_int, c_int]),
    ("imps_dev_copy", c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    ("imps_dev_alloc", c_int, [c_void_p, c_size_t, ctypes.POINTER(c_void_p)]),
    ("imps_dev_free", c_int, [c_void_p, c_void_p]),
    ("imps_memcpy_h2d", c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    ("imps_memcpy_d2h", c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    ("imps_synchronize", c_int, [c_void_p]),
]

_lib = None


def load():
    """Loads libimps.so and declares every C-ABI signature; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with infinitempsalgorithms_amd.build.build() "
                          "(there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


class ImpsError(RuntimeError):
    pass


def check(ctx_handle, status, what):
    if status != IMPS_OK:
        msg = load().imps_last_error(ctx_handle)
        raise ImpsError(f"{what} failed with status {status}: {msg.decode() if msg else ''}")


def dptr(a: np.ndarray):
    """Pointer to a float64 Fortran-ordered array (the caller keeps `a` alive)."""
    assert a.dtype == np.float64 and a.flags.f_contiguous
    return a.ctypes.data_as(P_double)


def ptr_array(arrays):
    """double*[] built from a list of Fortran-ordered float64 arrays."""
    arr = (P_double * len(arrays))()
    for i, a in enumerate(arrays):
        arr[i] = dptr(a)
    return ctypes.cast(arr, PP_double), arr


def fortran(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def zfortran(a):
    """complex128 Fortran-ordered copy (ComplexF64 layout: re, im interleaved)."""
    return np.asfortranarray(np.asarray(a, dtype=np.complex128))


def zptr(a: np.ndarray):
    """double* to a complex128 Fortran-ordered array (2 doubles per entry)."""
    assert a.dtype == np.complex128 and a.flags.f_contiguous
    return a.ctypes.data_as(P_double)


def zptr_array(arrays):
    """double*[] built from a list of Fortran-ordered complex128 arrays."""
    arr = (P_double * len(arrays))()
    for i, a in enumerate(arrays):
        arr[i] = zptr(a)
    return ctypes.cast(arr, PP_double), arr
