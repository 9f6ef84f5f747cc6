"""ctypes binding of libiqs_hip.so (the C ABI in include/iqs_hip.h).

The HIP library is the only compute path: if it is missing, or no GPU is visible, every
call raises.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libiqs_hip.so")

C64, C128 = 0, 1
HEA_GRAD, HEA_KEEP_STATE, HEA_NO_GRAPH = 1, 2, 4


class HeaTune(C.Structure):
    _fields_ = [("variant", C.c_int), ("workers_per_cu", C.c_int), ("order", C.c_int), ("group_mib", C.c_int)]


class HeaTrial(C.Structure):
    _fields_ = [("tune", HeaTune), ("ms", C.c_double)]

_lib = None


class IqsError(RuntimeError):
    pass


def _sig(L):
    P, U, U64, D, I = C.c_void_p, C.c_uint, C.c_uint64, C.c_double, C.c_int
    DP = C.POINTER(C.c_double)
    UP = C.POINTER(C.c_uint)
    sigs = {
        "iqsh_set_stream": [P],
        "iqsh_sync": [],
        "iqsh_device_cus": [C.POINTER(C.c_int)],
        "iqsh_timer_start": [],
        "iqsh_timer_stop": [DP],
        "iqsh_create": [U, I, C.POINTER(P)],
        "iqsh_wrap": [U, I, P, C.POINTER(P)],
        "iqsh_destroy": [P],
        "iqsh_info": [P, UP, C.POINTER(C.c_int), C.POINTER(P)],
        "iqsh_upload": [P, P, U64, U64],
        "iqsh_download": [P, P, U64, U64],
        "iqsh_copy": [P, P],
        "iqsh_convert": [P, P],
        "iqsh_init_basis": [P, U64],
        "iqsh_init_uniform": [P, D, D],
        "iqsh_apply_1q": [P, U, P],
        "iqsh_apply_1q_spec": [P, U, I, D],
        "iqsh_apply_nest": [P, P, P],
        "iqsh_spec_matrix": [I, I, I, D, P],
        "iqsh_apply_c1q": [P, U, U, P],
        "iqsh_apply_c1q_spec": [P, U, U, I, D],
        "iqsh_apply_mc1q": [P, U64, U, P],
        "iqsh_apply_swaplike": [P, U, U, P],
        "iqsh_apply_diag2": [P, U, U, P],
        "iqsh_apply_2q": [P, U, U, P],
        "iqsh_apply_kq": [P, U, P, P],
        "iqsh_set_block_mfma": [I],
        "iqsh_scale": [P, D, D],
        "iqsh_axpy": [P, P, D, D],
        "iqsh_collapse": [P, U, I],
        "iqsh_permute": [P, P],
        "iqsh_prob1": [P, U, DP],
        "iqsh_norm2": [P, DP],
        "iqsh_maxnorm_half": [P, U, I, DP],
        "iqsh_expval_pauli": [P, U64, U64, U, DP, DP],
        "iqsh_expval_pauli_noalloc": [P, U64, U64, U, DP, DP],
        "iqsh_overlap": [P, P, DP, DP],
        "iqsh_maxabsdiff": [P, P, D, D, DP],
        "iqsh_l2diff2": [P, P, DP],
        "iqsh_qaoa_phase": [P, U, P, P, P, D],
        "iqsh_qaoa_expect": [P, U, P, P, P, DP],
        "iqsh_maxcut_diag": [P, U, P, P, P, DP],
        "iqsh_qaoa_phase_diag": [P, P, D],
        "iqsh_qaoa_expect_diag": [P, P, I, DP],
        "iqsh_qaoa_cost_apply": [P, P, U, P, P, P, DP],
        "iqsh_qaoa_phase_adjoint": [P, P, U, P, P, P, D, DP],
        "iqsh_adjoint_1q": [P, P, U, P, P, DP],
        "iqsh_apply_1q_all": [P, P],
        "iqsh_adjoint_1q_all": [P, P, P, P, DP],
        "iqsh_qaoa_layer": [P, U, P, P, P, D, P],
        "iqsh_qaoa_layer_adjoint": [P, P, U, P, P, P, D, P, P, DP, DP],
        "iqsh_hea_variants": [I, C.POINTER(I)],
        "iqsh_hea_variant_info": [I, I, C.POINTER(I), C.POINTER(I), C.POINTER(I), C.POINTER(I)],
        "iqsh_hea_create": [I, I, P, P, P, C.POINTER(P)],
        "iqsh_hea_destroy": [P],
        "iqsh_hea_num_params": [P, C.POINTER(U64)],
        "iqsh_hea_set_tune": [P, P],
        "iqsh_hea_get_tune": [P, P],
        "iqsh_hea_set_angles": [P, P],
        "iqsh_hea_set_angles_device": [P, P],
        "iqsh_hea_run": [P, I],
        "iqsh_hea_results": [P, P, P],
        "iqsh_hea_results_device": [P, P, P],
        "iqsh_hea_set_wait_limit": [P, C.c_uint32, I],
        "iqsh_hea_state": [P, I, P],
        "iqsh_hea_traffic": [P, I, DP],
        "iqsh_hea_autotune": [P, I, I, P, I, C.POINTER(I), DP],
    }
    for name, args in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.iqsh_last_error.argtypes = []
    L.iqsh_last_error.restype = C.c_char_p


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        # One HIP runtime per process: PyTorch ships its own libamdhip64.so.7 under the same
        # soname as the system runtime this library links, and the first one loaded serves
        # every later user.  PyTorch's GPU set-up fails ("No HIP GPUs are available") once
        # this library has used the runtime first, so when PyTorch is installed and sees a
        # GPU, its set-up runs here, before the first call into this library.
        try:
            import torch
        except ImportError:
            torch = None
        if torch is not None and torch.cuda.is_available():
            torch.cuda.init()
        if not os.path.exists(LIB_PATH):
            raise IqsError(f"HIP engine library missing: {LIB_PATH} (build with `make -C intel-qs_amd`)")
        L = C.CDLL(LIB_PATH)
        _sig(L)
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        raise IqsError(f"iqs_hip error {rc}: {lib().iqsh_last_error().decode()}")


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))
