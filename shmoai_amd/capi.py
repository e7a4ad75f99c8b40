"""ctypes binding of libshmoai.so (the C ABI declared in include/shmoai.h).

The library is built in-tree by shmoai_amd/build.py. Loading fails loudly when
it is missing: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libshmoai.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "shmoai.h")

SHM_OK = 0
SHM_ERR_INVALID_ARGUMENT = -1
SHM_ERR_LOGIC = -2
SHM_ERR_RUNTIME = -3
SHM_ERR_OUT_OF_MEMORY = -4
SHM_ERR_DEVICE = -5

MATMUL_U64 = 0
MATMUL_PER_CT = 1
MATMUL_INT8 = 2

vp = C.c_void_p
u32 = C.c_uint32
u64 = C.c_uint64
sz = C.c_size_t
i32 = C.c_int
dbl = C.c_double
pu64 = C.POINTER(C.c_uint64)
pvp = C.POINTER(C.c_void_p)


class BootParams(C.Structure):
    _fields_ = [("s2c_input_primes", u32), ("layer3_factor", C.c_double), ("K", C.c_int32), ("deg", C.c_int32),
                ("double_angles", C.c_int32), ("input_bound", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", u64), ("completed", u64), ("failed", u64), ("ciphertexts", u64), ("groups", u64),
                ("in_flight_bytes", u64), ("p50_ms", dbl), ("p99_ms", dbl), ("busy_s", dbl),
                ("packed_requests", u64), ("shared_cts", u64), ("free_bytes", u64), ("admit_cap_bytes", u64),
                ("workers", u32)]


class ContextInfo(C.Structure):
    _fields_ = [("poly_degree", u32), ("size_q", u32), ("size_p", u32), ("dnum", u32)]


_SIGS = {
    "shm_last_error": (C.c_char_p, []),
    "shm_version": (C.c_char_p, []),
    "shm_context_create": (i32, [u32, C.POINTER(C.c_int), u32, u32, pvp]),
    "shm_context_destroy": (i32, [vp]),
    "shm_context_get_info": (i32, [vp, C.POINTER(ContextInfo)]),
    "shm_context_coeff_modulus": (i32, [vp, pu64, sz]),
    "shm_galois_elt_from_step": (i32, [vp, i32, C.POINTER(u32)]),
    "shm_synchronize": (i32, [vp]),
    "shm_ntt_forward": (i32, [vp, vp, u32, u32, u32, vp]),
    "shm_ntt_inverse": (i32, [vp, vp, u32, u32, u32, vp]),
    "shm_ciphertext_create": (i32, [vp, u32, u32, pvp]),
    "shm_ciphertext_destroy": (i32, [vp]),
    "shm_ciphertext_info": (i32, [vp, C.POINTER(u32), C.POINTER(u32), C.POINTER(u32), C.POINTER(dbl)]),
    "shm_ciphertext_set_scale": (i32, [vp, dbl]),
    "shm_ciphertext_data": (vp, [vp]),
    "shm_shared_a_batch_data": (vp, [vp]),
    "shm_ciphertext_import": (i32, [vp, pu64, sz, vp]),
    "shm_ciphertext_export": (i32, [vp, pu64, sz, vp]),
    "shm_ciphertext_copy": (i32, [vp, vp, vp]),
    "shm_plaintext_create": (i32, [vp, u32, pvp]),
    "shm_plaintext_destroy": (i32, [vp]),
    "shm_plaintext_info": (i32, [vp, C.POINTER(u32), C.POINTER(u32), C.POINTER(dbl)]),
    "shm_plaintext_set_scale": (i32, [vp, dbl]),
    "shm_plaintext_data": (vp, [vp]),
    "shm_plaintext_import": (i32, [vp, pu64, sz, vp]),
    "shm_plaintext_export": (i32, [vp, pu64, sz, vp]),
    "shm_secret_key_create": (i32, [vp, C.POINTER(C.c_int8), pvp]),
    "shm_secret_key_generate": (i32, [vp, u64, u32, pvp]),
    "shm_secret_key_destroy": (i32, [vp]),
    "shm_secret_key_export": (i32, [vp, pu64, sz]),
    "shm_relin_key_generate": (i32, [vp, vp, u64, pvp]),
    "shm_kswitch_key_generate": (i32, [vp, vp, pu64, u64, pvp]),
    "shm_kswitch_key_destroy": (i32, [vp]),
    "shm_kswitch_key_export": (i32, [vp, pu64, sz]),
    "shm_galois_keys_generate": (i32, [vp, vp, C.POINTER(C.c_int), sz, u64, pvp]),
    "shm_galois_keys_destroy": (i32, [vp]),
    "shm_galois_keys_export": (i32, [vp, u32, pu64, sz]),
    "shm_encrypt_symmetric": (i32, [vp, vp, vp, u64, vp, vp]),
    "shm_decrypt": (i32, [vp, vp, vp, vp, vp]),
    "shm_encode": (i32, [vp, C.POINTER(dbl), C.POINTER(dbl), sz, u32, dbl, vp, vp]),
    "shm_encode_scalar": (i32, [vp, dbl, u32, dbl, vp, vp]),
    "shm_decode": (i32, [vp, vp, C.POINTER(dbl), C.POINTER(dbl), sz, vp]),
    "shm_add_inplace": (i32, [vp, vp, vp, vp]),
    "shm_sub_inplace": (i32, [vp, vp, vp, vp]),
    "shm_negate_inplace": (i32, [vp, vp, vp]),
    "shm_add_plain_inplace": (i32, [vp, vp, vp, vp]),
    "shm_sub_plain_inplace": (i32, [vp, vp, vp, vp]),
    "shm_multiply_plain_inplace": (i32, [vp, vp, vp, vp]),
    "shm_multiply_inplace": (i32, [vp, vp, vp, vp]),
    "shm_square_inplace": (i32, [vp, vp, vp]),
    "shm_relinearize_inplace": (i32, [vp, vp, vp, vp]),
    "shm_apply_kswitch_key_inplace": (i32, [vp, vp, vp, vp]),
    "shm_rescale_to_next_inplace": (i32, [vp, vp, vp]),
    "shm_mod_switch_to_next_inplace": (i32, [vp, vp, vp]),
    "shm_mod_switch_to_inplace": (i32, [vp, vp, u32, vp]),
    "shm_plain_mod_switch_to_inplace": (i32, [vp, vp, u32, vp]),
    "shm_apply_galois_inplace": (i32, [vp, vp, u32, vp, vp]),
    "shm_rotate_vector_inplace": (i32, [vp, vp, i32, vp, vp]),
    "shm_complex_conjugate_inplace": (i32, [vp, vp, vp, vp]),
    "shm_hoisting_inplace": (i32, [vp, vp, vp, C.POINTER(C.c_int), sz, vp]),
    "shm_ct_pt_matrix_mul": (i32, [vp, pvp, u32, C.POINTER(dbl), u32, pvp, i32, vp]),
    "shm_ct_pt_matrix_mul_masked": (i32, [vp, pvp, u32, C.POINTER(dbl), u32, C.POINTER(C.c_int32), sz, pvp, vp]),
    "shm_last_matmul_kernel_ms": (dbl, []),
    "shm_last_matmul_rescale_tile_ms": (dbl, []),
    "shm_last_matmul_variant": (i32, []),
    "shm_bootstrapper_create": (i32, [vp, C.POINTER(BootParams), pvp]),
    "shm_bootstrapper_create_plan": (i32, [u32, u32, C.POINTER(BootParams), pvp]),
    "shm_bootstrapper_destroy": (i32, [vp]),
    "shm_bootstrapper_rotation_steps": (i32, [vp, C.POINTER(C.c_int), sz, C.POINTER(sz)]),
    "shm_bootstrapper_layer_info": (i32, [vp, i32, i32, C.POINTER(C.c_int64)]),
    "shm_bootstrapper_layer_diagonal": (i32, [vp, i32, i32, i32, i32, C.POINTER(dbl), C.POINTER(dbl), sz]),
    "shm_bootstrapper_evalmod_coeffs": (i32, [vp, C.POINTER(dbl), sz, C.POINTER(sz)]),
    "shm_bootstrapper_prepare": (i32, [vp, vp]),
    "shm_modraise_inplace": (i32, [vp, vp, vp]),
    "shm_multiply_by_i_inplace": (i32, [vp, vp, i32, vp]),
    "shm_bootstrap_slot_to_coeff": (i32, [vp, vp, vp, vp]),
    "shm_bootstrap_coeff_to_slot": (i32, [vp, vp, vp, vp, vp, vp]),
    "shm_bootstrap_eval_mod": (i32, [vp, vp, vp, vp]),
    "shm_bootstrap_eval_mod_batch": (i32, [vp, vp, sz, vp, vp]),
    "shm_bootstrap_s2c_first": (i32, [vp, pvp, sz, pvp, vp, vp, dbl, vp]),
    "shm_stream_create": (i32, [pvp]),
    "shm_stream_destroy": (i32, [vp]),
    "shm_graph_capture_begin": (i32, [vp]),
    "shm_graph_capture_end": (i32, [vp, pvp]),
    "shm_graph_launch": (i32, [vp, vp]),
    "shm_graph_destroy": (i32, [vp]),
    "shm_server_create": (i32, [vp, vp, vp, u64, u32, pvp]),
    "shm_server_destroy": (i32, [vp]),
    "shm_server_submit": (i32, [vp, pvp, sz, pvp, C.POINTER(u64)]),
    "shm_server_wait": (i32, [vp, u64, C.POINTER(dbl)]),
    "shm_server_get_stats": (i32, [vp, vp]),
    "shm_server_create_ex": (i32, [vp, vp, vp, u64, u32, u32, pvp]),
    "shm_server_submit_packed": (i32, [vp, pvp, sz, u32, pvp, C.POINTER(u64)]),
    "shm_server_wait_packed": (i32, [vp, u64, C.POINTER(dbl), C.POINTER(u32)]),
    "shm_evaluate_chebyshev": (i32, [vp, pvp, sz, C.POINTER(dbl), sz, dbl, dbl, vp, vp]),
    "shm_ct_ct_matrix_mul_colpacking": (i32, [vp, pvp, pvp, i32, i32, i32, i32, i32, vp, vp, pvp, vp]),
    "shm_ct_ct_matrix_mul_diagpacking": (i32, [vp, pvp, pvp, i32, i32, i32, i32, i32, vp, vp, pvp, vp]),
    "shm_shared_a_keys_create": (i32, [vp, vp, u32, u32, u64, u32, pvp]),
    "shm_shared_a_keys_destroy": (i32, [vp]),
    "shm_shared_a_keys_add_galois": (i32, [vp, C.POINTER(C.c_int), sz, u64]),
    "shm_shared_a_keys_aux_coeffs": (i32, [vp, u32, C.POINTER(C.c_int8), sz]),
    "shm_shared_a_keys_export_format": (i32, [vp, pu64, pu64, sz]),
    "shm_shared_a_keys_export_galois": (i32, [vp, u32, pu64, pu64, sz]),
    "shm_shared_a_batch_create": (i32, [vp, u32, u32, pvp]),
    "shm_shared_a_batch_destroy": (i32, [vp]),
    "shm_shared_a_batch_info": (i32, [vp, C.POINTER(u32), C.POINTER(u32), C.POINTER(u32), C.POINTER(dbl)]),
    "shm_shared_a_batch_set_scale": (i32, [vp, dbl]),
    "shm_shared_a_batch_import": (i32, [vp, pu64, sz, vp]),
    "shm_shared_a_batch_export": (i32, [vp, pu64, sz, vp]),
    "shm_shared_a_batch_member": (i32, [vp, u32, vp, vp]),
    "shm_fmt_switch_forward": (i32, [vp, pvp, sz, vp, vp]),
    "shm_fmt_switch_backward": (i32, [vp, vp, pvp, sz, vp]),
    "shm_shared_a_rotate_inplace": (i32, [vp, vp, i32, vp]),
    "shm_shared_a_add_inplace": (i32, [vp, vp, vp]),
    "shm_shared_a_multiply_plain_inplace": (i32, [vp, vp, vp]),
    "shm_shared_a_rescale_to_next_inplace": (i32, [vp, vp]),
    "shm_shared_a_modraise_inplace": (i32, [vp, vp]),
    "shm_bootstrap_shared_a_s2c_first": (i32, [vp, vp, pvp, sz, pvp, vp, dbl, vp]),
    "shm_bootstrap_shared_slot_to_coeff": (i32, [vp, vp, vp, vp]),
    "shm_bootstrap_shared_coeff_to_slot": (i32, [vp, vp, vp, vp, vp, vp]),
    "shm_public_key_generate": (i32, [vp, vp, u64, pvp]),
    "shm_public_key_destroy": (i32, [vp]),
    "shm_public_key_export": (i32, [vp, pu64, sz]),
    "shm_encrypt_asymmetric": (i32, [vp, vp, vp, u64, vp, vp]),
    "shm_batch_input": (i32, [vp, vp, C.POINTER(dbl), i32, i32, i32, dbl, u64, pvp, vp]),
    "shm_bias_vec": (i32, [C.POINTER(C.c_int), i32, i32, C.POINTER(C.c_int), sz]),
    "shm_multiply_const_inplace": (i32, [vp, vp, dbl, vp]),
    "shm_multiply_const_complex_inplace": (i32, [vp, vp, dbl, dbl, vp]),
    "shm_add_const_inplace": (i32, [vp, vp, dbl, vp]),
    "shm_multiply_vector_inplace_reduced_error": (i32, [vp, vp, C.POINTER(dbl), C.POINTER(dbl), sz, vp]),
    "shm_add_inplace_reduced_error": (i32, [vp, vp, vp, vp]),
    "shm_sub_inplace_reduced_error": (i32, [vp, vp, vp, vp]),
    "shm_multiply_inplace_reduced_error": (i32, [vp, vp, vp, vp, vp]),
    "shm_double_inplace": (i32, [vp, vp, vp]),
    "shm_rescale_to_next_inplace_noalloc": (i32, [vp, vp, vp]),
    "shm_bootstrap_shared_a_s2c_first_chunks": (i32, [vp, vp, pvp, sz, pvp, vp, dbl, vp]),
    "shm_boot_graph_create_shared_a": (i32, [vp, vp, vp, dbl, dbl, vp, pvp]),
    "shm_boot_graph_create_ordinary": (i32, [vp, vp, vp, u32, dbl, dbl, vp, pvp]),
    "shm_boot_graph_run": (i32, [vp, pvp, sz, pvp]),
    "shm_boot_graph_info": (i32, [vp, C.POINTER(u32), C.POINTER(u64)]),
    "shm_boot_graph_destroy": (i32, [vp]),
    "shm_probe_fused_ntt_begin": (i32, []),
    "shm_probe_fused_ntt_end": (i32, [C.POINTER(u64), C.POINTER(dbl), C.POINTER(dbl)]),
    "shm_probe_plain_ntt_last": (i32, [C.POINTER(u64), C.POINTER(dbl), C.POINTER(dbl)]),
    "shm_keyswitch_counters": (i32, [pu64, pu64, pu64, sz]),
    "shm_keyswitch_counters_reset": (i32, []),
    "shm_set_tuning": (i32, [C.c_char_p, i32]),
    "shm_get_tuning": (i32, [C.c_char_p, C.POINTER(i32)]),
    "shm_autotune_enable": (i32, [i32]),
    "shm_autotune_clear": (i32, []),
    "shm_autotune_report": (i32, [C.c_char_p, sz, C.POINTER(sz)]),
}

_lib = None


class ShmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class ShmInvalidArgument(ShmError, ValueError):
    pass


class ShmLogicError(ShmError):
    pass


_ERRS = {SHM_ERR_INVALID_ARGUMENT: ShmInvalidArgument, SHM_ERR_LOGIC: ShmLogicError}


def header_functions() -> list[str]:
    """Every function name declared in include/shmoai.h."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(shm_[a-z0-9_]+)\s*\(", txt)))


def load(path: str = LIB_PATH):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with `python -m shmoai_amd.build` "
                          "(there is no fallback implementation)")
    L = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != SHM_OK:
        msg = load().shm_last_error().decode(errors="replace")
        raise _ERRS.get(rc, ShmError)(rc, msg)
