"""ctypes binding of libcufcm_fil.so (declarations in include/cufcm_fil.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcufcm_fil.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "cufcm_fil.h")
# measurement helper beside the mobility library (include/cfm_probe.h)
PROBE_LIB_PATH = os.path.join(HERE, "libcfm_probe.so")
PROBE_HEADER = os.path.join(os.path.dirname(HERE), "include", "cfm_probe.h")

CFM_OK, CFM_EINVAL, CFM_EHIP, CFM_ENOMEM, CFM_ESTATE, CFM_ENOCONV, CFM_ENONFINITE = range(7)
CFM_FP64, CFM_FP32C = 0, 1
CFM_SS, CFM_SB, CFM_BB, CFM_BS, CFM_FULL = range(5)
CFM_SYS_SEG, CFM_SYS_BLOB, CFM_SYS_BODY = 0, 1, 2
CFM_PC_NONE, CFM_PC_DIAG, CFM_PC_SELF, CFM_PC_NEAR, CFM_PC_DENSE = 0, 1, 2, 3, 4
CFM_SOLVE_FP64, CFM_SOLVE_MIXED = 0, 1
CFM_COARSE_NONE, CFM_COARSE_LINEAR, CFM_COARSE_LEARNED = 0, 1, 2
CFM_BODY_FREE, CFM_BODY_PRESCRIBED_VELOCITY, CFM_BODY_PRESCRIBED_ROTATION = 0, 1, 2


class CfmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[cfm error {code}] {msg}")
        self.code = code


class cfm_config(ctypes.Structure):
    _fields_ = [
        ("n_seg", ctypes.c_int32),
        ("n_blob", ctypes.c_int32),
        ("seg_dof", ctypes.c_int32),
        ("wall", ctypes.c_int32),
        ("a_seg", ctypes.c_double),
        ("a_blob", ctypes.c_double),
        ("mu", ctypes.c_double),
        ("precision", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("tile_split", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class cfm_solve_opts(ctypes.Structure):
    _fields_ = [
        ("system", ctypes.c_int32),
        ("precond", ctypes.c_int32),
        ("mode", ctypes.c_int32),
        ("restart", ctypes.c_int32),
        ("max_iter", ctypes.c_int32),
        ("use_guess", ctypes.c_int32),
        ("rtol", ctypes.c_double),
    ]


class cfm_solve_stats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int32),
        ("refinements", ctypes.c_int32),
        ("converged", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("rel_residual", ctypes.c_double),
        ("initial_rel_residual", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


CFM_GUESS_PREVIOUS, CFM_GUESS_LEARNED, CFM_GUESS_PROJECTED = 0, 1, 2


class cfm_fil_params(ctypes.Structure):
    _fields_ = [
        ("nseg", ctypes.c_int32),
        ("max_broyden", ctypes.c_int32),
        ("num_euler_steps", ctypes.c_int32),
        ("barrier", ctypes.c_int32),
        ("dl", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("kb", ctypes.c_double),
        ("kt", ctypes.c_double),
        ("end_force", ctypes.c_double),
        ("repulsive_factor", ctypes.c_double),
        ("tol", ctypes.c_double),
        ("gamma", ctypes.c_double),
        ("pair_precision", ctypes.c_int32),
        ("guess", ctypes.c_int32),
        ("blob_solve", cfm_solve_opts),
        ("overlap", ctypes.c_int32),
        ("broyden_only", ctypes.c_int32),
    ]


class cfm_step_stats(ctypes.Structure):
    _fields_ = [
        ("broyden_iterations", ctypes.c_int32),
        ("gmres_iterations", ctypes.c_int32),
        ("gmres_solves", ctypes.c_int32),
        ("converged", ctypes.c_int32),
        ("final_error", ctypes.c_double),
        ("max_gmres_rel_residual", ctypes.c_double),
        ("seconds", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class cfm_profile_stats(ctypes.Structure):
    _fields_ = [
        ("launches", ctypes.c_int64),
        ("kernel_ms", ctypes.c_double),
        ("alg_flops", ctypes.c_double),
        ("alg_bytes", ctypes.c_double),
        ("launches_fp32", ctypes.c_int64),
        ("kernel_ms_fp32", ctypes.c_double),
        ("alg_flops_fp32", ctypes.c_double),
        ("launches_sym", ctypes.c_int64),
        ("kernel_ms_sym", ctypes.c_double),
        ("alg_flops_sym", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


CFM_BEAT_FULFORD_BLAKE = 1
CFM_BEAT_FULFORD_BLAKE_ORIGINAL = 2
CFM_BEAT_VOLVOX = 3
CFM_BEAT_CORAL_LARVAE = 4
CFM_BEAT_BUILD_A_BEAT = 5
CFM_BEAT_BICILIA = 6
CFM_BEAT_BICILIA_LONGT = 7


class cfm_cilia_params(ctypes.Structure):
    _fields_ = [
        ("n_cilia", ctypes.c_int32),
        ("nseg", ctypes.c_int32),
        ("beat", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("dl", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("s_tol", ctypes.c_double),
        ("solve", cfm_solve_opts),
    ]


class cfm_squirmer_params(ctypes.Structure):
    _fields_ = [
        ("n_modes", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("A", ctypes.c_double * 8),
        ("B", ctypes.c_double * 8),
        ("dt", ctypes.c_double),
        ("solve", cfm_solve_opts),
    ]


class cfm_batch_config(ctypes.Structure):
    _fields_ = [
        ("n_sys", ctypes.c_int32),
        ("seg_dof", ctypes.c_int32),
        ("wall", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("a_seg", ctypes.c_double),
        ("mu", ctypes.c_double),
        ("reserved", ctypes.c_int32 * 4),
    ]


class cfm_batch_tuning(ctypes.Structure):
    _fields_ = [
        ("max_sources", ctypes.c_int32),
        ("grid", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 2),
        ("ms", ctypes.c_double),
    ]

    def as_dict(self):
        return {"max_sources": self.max_sources, "grid": self.grid, "ms": self.ms}


class cfm_batch_sys_stats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int32),
        ("converged", ctypes.c_int32),
        ("reserved0", ctypes.c_int32),
        ("reserved1", ctypes.c_int32),
        ("rel_residual", ctypes.c_double),
        ("latency_ms", ctypes.c_double),
    ]


class cfm_cilia_batch_config(ctypes.Structure):
    _fields_ = [
        ("n_members", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("spring_model", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("mu", ctypes.c_double),
        ("spring", ctypes.c_double),
        ("noise", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("prm", cfm_cilia_params),
    ]


_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_lib = None


def header_symbols(header: str = HEADER) -> list[str]:
    """Function names declared in a header (default include/cufcm_fil.h)."""
    with open(header) as fh:
        text = fh.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cfm_[a-z0-9_]+)\s*\(", text)))


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -m cufcm_fil_amd.build` "
            "(there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    L.cfm_abi_version.restype = ctypes.c_int
    L.cfm_last_error.restype = ctypes.c_char_p
    L.cfm_last_error.argtypes = [_vp]
    L.cfm_create.argtypes = [ctypes.POINTER(cfm_config), ctypes.POINTER(_vp)]
    L.cfm_destroy.argtypes = [_vp]
    L.cfm_set_positions.argtypes = [_vp, _dp, _dp]
    L.cfm_set_precision.argtypes = [_vp, ctypes.c_int32]
    L.cfm_apply.argtypes = [_vp, ctypes.c_int32, _dp, _dp, _dp, _dp]
    L.cfm_time_apply.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, _dp]
    L.cfm_apply_noalloc.argtypes = [_vp, ctypes.c_int32, _vp, _vp, _vp, _vp]
    L.cfm_gmres.argtypes = [_vp, ctypes.POINTER(cfm_solve_opts), _dp, _dp,
                            ctypes.POINTER(cfm_solve_stats)]
    L.cfm_apply_preconditioner.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, _dp, _dp]
    L.cfm_system_apply.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, _dp, _dp]
    L.cfm_body_init.argtypes = [_vp, ctypes.c_int32, _dp]
    L.cfm_body_set_state.argtypes = [_vp, _dp, _dp]
    L.cfm_body_reference_matrices.argtypes = [_vp, ctypes.c_int32, ctypes.POINTER(cfm_solve_opts),
                                              _dp, _dp, ctypes.POINTER(cfm_solve_stats)]
    L.cfm_body_set_reference_matrices.argtypes = [_vp, _dp, _dp]
    L.cfm_body_get_reference_matrices.argtypes = [_vp, _dp, _dp]
    L.cfm_body_set_motion.argtypes = [_vp, ctypes.c_int32]
    L.cfm_body_get_motion.argtypes = [_vp, ctypes.POINTER(ctypes.c_int32)]
    L.cfm_fil_init.argtypes = [_vp, ctypes.POINTER(cfm_fil_params), _dp, _dp]
    L.cfm_fil_step.argtypes = [_vp, ctypes.c_int32, ctypes.POINTER(cfm_step_stats)]
    L.cfm_fil_get_state.argtypes = [_vp, _dp, _dp, _dp, _dp, _dp]
    L.cfm_fil_set_state.argtypes = [_vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    L.cfm_fil_internal_forces.argtypes = [_vp, ctypes.c_int32, _dp]
    L.cfm_barrier_forces.argtypes = [_vp, _dp]
    L.cfm_fil_jacobian_inverse.argtypes = [_vp, ctypes.c_int32, _dp]
    L.cfm_fil_errors.argtypes = [_vp, ctypes.c_int32, _dp, _dp, _dp]
    L.cfm_fil_update.argtypes = [_vp, _dp]
    L.cfm_fil_initial_guess.argtypes = [_vp, ctypes.c_int32]
    L.cfm_fil_attach_bodies.argtypes = [_vp, _dp]
    L.cfm_fil_get_body_state.argtypes = [_vp, _dp, _dp, _dp, _dp]
    L.cfm_fil_get_blob_forces.argtypes = [_vp, _dp]
    L.cfm_fil_set_body_sphere.argtypes = [_vp, ctypes.c_double]
    _fp = ctypes.POINTER(ctypes.c_float)
    _ip = ctypes.POINTER(ctypes.c_int32)
    L.cfm_gnn_init.argtypes = [_vp, ctypes.c_uint32, ctypes.c_int32]
    L.cfm_gnn_num_params.argtypes = [_vp]
    L.cfm_gnn_record.argtypes = [_vp, ctypes.c_int32]
    L.cfm_gnn_samples.argtypes = [_vp]
    L.cfm_gnn_add_samples.argtypes = [_vp, ctypes.c_int32, _dp, _dp]
    L.cfm_gnn_train.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_double, _dp, _dp]
    L.cfm_gnn_enable.argtypes = [_vp, ctypes.c_int32]
    L.cfm_gnn_get_params.argtypes = [_vp, _fp]
    L.cfm_gnn_set_params.argtypes = [_vp, _fp]
    L.cfm_gnn_correction.argtypes = [_vp, _dp, _ip]
    L.cfm_gnn_gradient.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, _fp, _dp]
    L.cfm_gemm_bf16.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _fp,
                                ctypes.c_int32, _fp, ctypes.c_int32, _fp]
    L.cfm_nearfield_coarse.argtypes = [_vp, ctypes.c_int32]
    L.cfm_squirmer_init.argtypes = [_vp, ctypes.POINTER(cfm_squirmer_params)]
    L.cfm_squirmer_step.argtypes = [_vp, ctypes.c_int32, _dp, ctypes.POINTER(cfm_solve_stats)]
    L.cfm_gnn_train_coarse.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_double, _dp, _dp]
    L.cfm_gnn_coarse_gradient.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, _fp, _dp, _dp]
    L.cfm_gnn_coarse_vectors.argtypes = [_vp, _dp]
    L.cfm_fil_set_members.argtypes = [_vp, ctypes.c_int32]
    L.cfm_fil_member_stats.argtypes = [_vp, ctypes.POINTER(cfm_step_stats)]
    L.cfm_cilia_init.argtypes = [_vp, ctypes.POINTER(cfm_cilia_params), _dp, _dp, _dp]
    L.cfm_cilia_step.argtypes = [_vp, ctypes.c_int32, _dp, _dp, ctypes.POINTER(cfm_solve_stats)]
    L.cfm_cilia_get_state.argtypes = [_vp, _dp, _dp, _dp, _dp, _dp, _dp]
    L.cfm_cilia_set_driving.argtypes = [_vp, _dp, ctypes.c_int32, _dp, ctypes.c_int32,
                                        ctypes.c_double]
    L.cfm_cilia_get_speeds.argtypes = [_vp, _dp, _dp, _dp, _dp]
    L.cfm_cilia_set_s_table.argtypes = [_vp, _dp, ctypes.c_int32, ctypes.c_int32]
    L.cfm_cilia_set_phases.argtypes = [_vp, _dp, _dp]
    L.cfm_cilia_set_pair_dp.argtypes = [_vp, ctypes.c_double]
    L.cfm_cilia_init_bodies.argtypes = [_vp, ctypes.POINTER(cfm_cilia_params), _dp, _dp, _dp, _dp]
    L.cfm_cilia_get_bodies.argtypes = [_vp, _dp, _dp, _dp, _dp]
    L.cfm_cilia_set_driving_options.argtypes = [_vp, ctypes.c_int32, ctypes.c_double, ctypes.c_uint64]
    L.cfm_batch_create.argtypes = [ctypes.POINTER(cfm_batch_config), _ip, ctypes.POINTER(_vp)]
    L.cfm_batch_destroy.argtypes = [_vp]
    L.cfm_batch_last_error.restype = ctypes.c_char_p
    L.cfm_batch_last_error.argtypes = [_vp]
    L.cfm_batch_set_positions.argtypes = [_vp, _dp]
    L.cfm_batch_apply.argtypes = [_vp, ctypes.c_int32, _dp, _dp]
    L.cfm_batch_set_tuning.argtypes = [_vp, ctypes.POINTER(cfm_batch_tuning)]
    L.cfm_batch_get_tuning.argtypes = [_vp, ctypes.POINTER(cfm_batch_tuning)]
    L.cfm_batch_tune.argtypes = [_vp, ctypes.c_int32, ctypes.POINTER(cfm_batch_tuning),
                                 ctypes.POINTER(cfm_batch_tuning), ctypes.c_int32, _ip]
    L.cfm_batch_gmres.argtypes = [_vp, _dp, _dp, ctypes.POINTER(cfm_solve_opts),
                                  ctypes.POINTER(cfm_batch_sys_stats), _dp]
    L.cfm_profile.argtypes = [_vp, ctypes.c_int32]
    L.cfm_profile_read.argtypes = [_vp, ctypes.POINTER(cfm_profile_stats)]
    L.cfm_set_pair_symmetry.argtypes = [_vp, ctypes.c_int32]
    L.cfm_cilia_batch_set_pair_dp.argtypes = [_vp, ctypes.c_double]
    L.cfm_cilia_batch_create.argtypes = [ctypes.POINTER(cfm_cilia_batch_config), _dp, _dp, ctypes.c_int32,
                                         _dp, ctypes.c_int32, ctypes.POINTER(_vp)]
    L.cfm_cilia_batch_destroy.argtypes = [_vp]
    L.cfm_cilia_batch_last_error.argtypes = [_vp]
    L.cfm_cilia_batch_last_error.restype = ctypes.c_char_p
    L.cfm_cilia_batch_set_state.argtypes = [_vp, _dp, _dp, _dp, _dp]
    L.cfm_cilia_batch_step.argtypes = [_vp, ctypes.c_int32, _dp, _dp, ctypes.POINTER(cfm_batch_sys_stats)]
    L.cfm_cilia_batch_get_state.argtypes = [_vp, _dp, _dp, _dp, _dp]
    L.cfm_cilia_batch_get_positions.argtypes = [_vp, _dp]
    L.cfm_cilia_batch_get_solution.argtypes = [_vp, _dp]
    _lib = L
    return L


_probe = None


def probe_lib():
    """libcfm_probe.so (the FP32 issue-rate probe; not part of the mobility library)."""
    global _probe
    if _probe is None:
        if not os.path.exists(PROBE_LIB_PATH):
            raise ImportError(f"{PROBE_LIB_PATH} is not built; run `python -m cufcm_fil_amd.build`")
        P = ctypes.CDLL(PROBE_LIB_PATH)
        P.cfm_probe_valu_rate.argtypes = [ctypes.c_int32, ctypes.c_int32, _dp]
        _probe = P
    return _probe


def ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


def check(ctx, code):
    if code != CFM_OK:
        msg = lib().cfm_last_error(ctx)
        raise CfmError(code, msg.decode() if msg else "")


def check_batch(b, code):
    if code != CFM_OK:
        msg = lib().cfm_batch_last_error(b)
        raise CfmError(code, msg.decode() if msg else "")


def check_cilia_batch(b, code):
    if code != CFM_OK:
        msg = lib().cfm_cilia_batch_last_error(b)
        raise CfmError(code, msg.decode() if msg else "")


def fptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def iptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
