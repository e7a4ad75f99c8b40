"""shmoai_amd — MI355X-native CKKS engine for ShMOAI's encrypted-BERT hot path.

Python mirror of the reference's evaluator API (the `moai::` facade in
src/include/source/ckks_evaluator_parallel.cuh and the PhantomFHE free
functions in thirdparty/phantom-fhe/include/evaluate.cuh), layered over the C
ABI of libshmoai.so (include/shmoai.h). Every operation runs in hand-written
HIP kernels for gfx950; nothing here computes on the host.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import capi
from .capi import check, MATMUL_U64, MATMUL_PER_CT, MATMUL_INT8  # noqa: F401

__all__ = [
    "Context", "Ciphertext", "Plaintext", "SecretKey", "KSwitchKey", "GaloisKeys", "Encoder", "Evaluator",
    "Decryptor", "ct_pt_matrix_mul_wo_pre", "MATMUL_U64", "MATMUL_PER_CT", "MATMUL_INT8", "Bootstrapper",
    "SharedAKeys", "SharedABatch", "PublicKey", "Encryptor", "batch_input", "bias_vec",
]


def _lib():
    return capi.load()


def _u64p(a: np.ndarray):
    return a.ctypes.data_as(capi.pu64)


class Context:
    """EncryptionParameters + CoeffModulus::Create + PhantomContext (context.cu:125).

    bit_sizes lists the data primes then the special primes, as passed to
    CoeffModulus::Create; special_modulus_size is set_special_modulus_size.
    """

    def __init__(self, log_n: int, bit_sizes: Sequence[int], special_modulus_size: int):
        L = _lib()
        arr = (C.c_int * len(bit_sizes))(*bit_sizes)
        h = C.c_void_p()
        check(L.shm_context_create(log_n, arr, len(bit_sizes), special_modulus_size, C.byref(h)))
        self.h = h
        info = capi.ContextInfo()
        check(L.shm_context_get_info(self.h, C.byref(info)))
        self.N = info.poly_degree
        self.size_q = info.size_q
        self.size_p = info.size_p
        self.dnum = info.dnum
        self.coeff_modulus = np.zeros(self.size_q + self.size_p, np.uint64)
        check(L.shm_context_coeff_modulus(self.h, _u64p(self.coeff_modulus), self.coeff_modulus.size))

    def close(self):
        if getattr(self, "h", None):
            _lib().shm_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def ql(self, chain_index: int) -> int:
        return self.size_q - chain_index + 1

    def chain(self, ql: int) -> int:
        return self.size_q - ql + 1

    def galois_elt_from_step(self, step: int) -> int:
        e = C.c_uint32()
        check(_lib().shm_galois_elt_from_step(self.h, step, C.byref(e)))
        return e.value

    # raw limb transforms on device memory (e.g. a torch tensor's data_ptr())
    def ntt_forward(self, dev_ptr: int, n_polys: int, n_limbs: int, first_prime: int = 0, stream=None):
        check(_lib().shm_ntt_forward(self.h, C.c_void_p(dev_ptr), n_polys, n_limbs, first_prime, stream))

    def ntt_inverse(self, dev_ptr: int, n_polys: int, n_limbs: int, first_prime: int = 0, stream=None):
        check(_lib().shm_ntt_inverse(self.h, C.c_void_p(dev_ptr), n_polys, n_limbs, first_prime, stream))


class Ciphertext:
    """PhantomCiphertext (ciphertext.h:21-288): [size][Ql][N] uint64 in NTT form."""

    def __init__(self, ctx: Context, size: int = 2, chain_index: int = 1):
        self.ctx = ctx
        h = C.c_void_p()
        check(_lib().shm_ciphertext_create(ctx.h, size, chain_index, C.byref(h)))
        self.h = h

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _lib().shm_ciphertext_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def info(self):
        s, ch, q = C.c_uint32(), C.c_uint32(), C.c_uint32()
        sc = C.c_double()
        check(_lib().shm_ciphertext_info(self.h, C.byref(s), C.byref(ch), C.byref(q), C.byref(sc)))
        return s.value, ch.value, q.value, sc.value

    @property
    def size(self):
        return self.info()[0]

    @property
    def chain_index(self):
        return self.info()[1]

    @property
    def coeff_modulus_size(self):
        return self.info()[2]

    @property
    def scale(self):
        return self.info()[3]

    @scale.setter
    def scale(self, v: float):
        check(_lib().shm_ciphertext_set_scale(self.h, float(v)))

    def data_ptr(self) -> int:
        return _lib().shm_ciphertext_data(self.h)

    def words(self) -> int:
        s, _, q, _ = self.info()
        return s * q * self.ctx.N

    def load(self, limbs: np.ndarray, stream=None) -> "Ciphertext":
        a = np.ascontiguousarray(limbs, np.uint64).ravel()
        check(_lib().shm_ciphertext_import(self.h, _u64p(a), a.size, stream))
        return self

    def numpy(self, stream=None) -> np.ndarray:
        out = np.zeros(self.words(), np.uint64)
        check(_lib().shm_ciphertext_export(self.h, _u64p(out), out.size, stream))
        return out

    def copy(self, stream=None) -> "Ciphertext":
        s, ch, _, _ = self.info()
        out = Ciphertext(self.ctx, s, ch)
        check(_lib().shm_ciphertext_copy(self.h, out.h, stream))
        return out


class Plaintext:
    """PhantomPlaintext: [Ql][N] uint64 in NTT form."""

    def __init__(self, ctx: Context, chain_index: int = 1):
        self.ctx = ctx
        h = C.c_void_p()
        check(_lib().shm_plaintext_create(ctx.h, chain_index, C.byref(h)))
        self.h = h

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _lib().shm_plaintext_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def info(self):
        ch, q = C.c_uint32(), C.c_uint32()
        sc = C.c_double()
        check(_lib().shm_plaintext_info(self.h, C.byref(ch), C.byref(q), C.byref(sc)))
        return ch.value, q.value, sc.value

    @property
    def chain_index(self):
        return self.info()[0]

    @property
    def scale(self):
        return self.info()[2]

    @scale.setter
    def scale(self, v: float):
        check(_lib().shm_plaintext_set_scale(self.h, float(v)))

    def load(self, limbs: np.ndarray, stream=None) -> "Plaintext":
        a = np.ascontiguousarray(limbs, np.uint64).ravel()
        check(_lib().shm_plaintext_import(self.h, _u64p(a), a.size, stream))
        return self

    def numpy(self, stream=None) -> np.ndarray:
        _, q, _ = self.info()
        out = np.zeros(q * self.ctx.N, np.uint64)
        check(_lib().shm_plaintext_export(self.h, _u64p(out), out.size, stream))
        return out


class KSwitchKey:
    """PhantomRelinKey / one Galois key: dnum x [2][QP][N]."""

    def __init__(self, ctx: Context, h):
        self.ctx, self.h = ctx, h

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _lib().shm_kswitch_key_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def numpy(self) -> np.ndarray:
        c = self.ctx
        out = np.zeros(c.dnum * 2 * (c.size_q + c.size_p) * c.N, np.uint64)
        check(_lib().shm_kswitch_key_export(self.h, _u64p(out), out.size))
        return out


class GaloisKeys:
    """PhantomGaloisKey for a set of rotation steps."""

    def __init__(self, ctx: Context, h, steps):
        self.ctx, self.h, self.steps = ctx, h, list(steps)

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _lib().shm_galois_keys_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def numpy(self, elt: int) -> np.ndarray:
        c = self.ctx
        out = np.zeros(c.dnum * 2 * (c.size_q + c.size_p) * c.N, np.uint64)
        check(_lib().shm_galois_keys_export(self.h, elt, _u64p(out), out.size))
        return out


class SecretKey:
    """PhantomSecretKey (secretkey.cu:431-646) with injected, seeded randomness."""

    def __init__(self, ctx: Context, coeffs: np.ndarray | None = None, seed: int = 0, hamming_weight: int = 64):
        self.ctx = ctx
        h = C.c_void_p()
        if coeffs is not None:
            a = np.ascontiguousarray(coeffs, np.int8)
            check(_lib().shm_secret_key_create(ctx.h, a.ctypes.data_as(C.POINTER(C.c_int8)), C.byref(h)))
        else:
            check(_lib().shm_secret_key_generate(ctx.h, seed, hamming_weight, C.byref(h)))
        self.h = h

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _lib().shm_secret_key_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def numpy(self) -> np.ndarray:
        c = self.ctx
        out = np.zeros((c.size_q + c.size_p) * c.N, np.uint64)
        check(_lib().shm_secret_key_export(self.h, _u64p(out), out.size))
        return out

    def gen_relinkey(self, seed: int) -> KSwitchKey:
        h = C.c_void_p()
        check(_lib().shm_relin_key_generate(self.ctx.h, self.h, seed, C.byref(h)))
        return KSwitchKey(self.ctx, h)

    def gen_kswitch_key(self, new_key_qp: np.ndarray, seed: int) -> KSwitchKey:
        a = np.ascontiguousarray(new_key_qp, np.uint64)
        h = C.c_void_p()
        check(_lib().shm_kswitch_key_generate(self.ctx.h, self.h, _u64p(a), seed, C.byref(h)))
        return KSwitchKey(self.ctx, h)

    def create_galois_keys_from_steps(self, steps: Sequence[int], seed: int) -> GaloisKeys:
        arr = (C.c_int * len(steps))(*steps)
        h = C.c_void_p()
        check(_lib().shm_galois_keys_generate(self.ctx.h, self.h, arr, len(steps), seed, C.byref(h)))
        return GaloisKeys(self.ctx, h, steps)

    def gen_publickey(self, seed: int) -> "PublicKey":
        h = C.c_void_p()
        check(_lib().shm_public_key_generate(self.ctx.h, self.h, seed, C.byref(h)))
        return PublicKey(self.ctx, h)

    def encrypt_symmetric(self, pt: Plaintext, seed: int, stream=None) -> Ciphertext:
        ct = Ciphertext(self.ctx, 2, pt.chain_index)
        check(_lib().shm_encrypt_symmetric(self.ctx.h, self.h, pt.h, seed, ct.h, stream))
        return ct


class PublicKey:
    """PhantomPublicKey (secretkey.h): (-(a s + e), a) over QP, NTT form."""

    def __init__(self, ctx: Context, h):
        self.ctx, self.h = ctx, h

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _lib().shm_public_key_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def numpy(self) -> np.ndarray:
        c = self.ctx
        out = np.zeros(2 * (c.size_q + c.size_p) * c.N, np.uint64)
        check(_lib().shm_public_key_export(self.h, _u64p(out), out.size))
        return out


class Encryptor:
    """moai::Encryptor (ckks_evaluator_parallel.cuh:107-127): public-key encryption."""

    def __init__(self, ctx: Context, pk: PublicKey):
        self.ctx, self.pk = ctx, pk

    def encrypt(self, pt: "Plaintext", seed: int, stream=None) -> "Ciphertext":
        ct = Ciphertext(self.ctx, 2, 1)
        check(_lib().shm_encrypt_asymmetric(self.ctx.h, self.pk.h, pt.h, seed, ct.h, stream))
        return ct


def batch_input(X: np.ndarray, scale: float, ctx: Context, pk: PublicKey, seed: int, stream=None) -> list:
    """batch_input (Batch_encode_encrypt.cuh:12-43): X[num_X][num_row][num_col] ->
    num_col ciphertexts, column i packed as slot num_X * k + j = X[j][k][i]."""
    X = np.ascontiguousarray(X, np.float64)
    num_x, num_row, num_col = X.shape
    out = [Ciphertext(ctx, 2, 1) for _ in range(num_col)]
    hs = (C.c_void_p * max(num_col, 1))(*[o.h.value for o in out])
    check(_lib().shm_batch_input(ctx.h, pk.h, X.ctypes.data_as(C.POINTER(C.c_double)), num_x, num_row, num_col,
                                 float(scale), seed, hs, stream))
    return out


def bias_vec(lengths: Sequence[int], num_x: int, num_row: int) -> np.ndarray:
    """bias_vec (Batch_encode_encrypt.cuh:45-57): token mask, out[j * num_X + i] = j < lengths[i]."""
    arr = (C.c_int * len(lengths))(*lengths)
    out = np.zeros(num_x * num_row, np.int32)
    check(_lib().shm_bias_vec(arr, num_x, num_row, out.ctypes.data_as(C.POINTER(C.c_int)), out.size))
    return out


class Decryptor:
    """moai::Decryptor (ckks_evaluator_parallel.cuh:564-598)."""

    def __init__(self, ctx: Context, sk: SecretKey):
        self.ctx, self.sk = ctx, sk

    def decrypt(self, ct: Ciphertext, stream=None) -> Plaintext:
        pt = Plaintext(self.ctx, ct.chain_index)
        check(_lib().shm_decrypt(self.ctx.h, self.sk.h, ct.h, pt.h, stream))
        return pt


class Encoder:
    """moai::Encoder over PhantomCKKSEncoder (ckks.cu:89-246)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx

    def slot_count(self) -> int:
        return self.ctx.N // 2

    def encode(self, values, scale: float, chain_index: int = 1, stream=None) -> Plaintext:
        pt = Plaintext(self.ctx, chain_index)
        if np.isscalar(values):
            check(_lib().shm_encode_scalar(self.ctx.h, float(values), chain_index, float(scale), pt.h, stream))
            return pt
        v = np.asarray(values)
        re = np.ascontiguousarray(v.real, np.float64)
        im = np.ascontiguousarray(v.imag, np.float64) if np.iscomplexobj(v) else None
        check(_lib().shm_encode(self.ctx.h, re.ctypes.data_as(C.POINTER(C.c_double)),
                                im.ctypes.data_as(C.POINTER(C.c_double)) if im is not None else None,
                                re.size, chain_index, float(scale), pt.h, stream))
        return pt

    def decode(self, pt: Plaintext, stream=None, complex_out: bool = False):
        n = self.ctx.N // 2
        re = np.zeros(n, np.float64)
        im = np.zeros(n, np.float64)
        check(_lib().shm_decode(self.ctx.h, pt.h, re.ctypes.data_as(C.POINTER(C.c_double)),
                                im.ctypes.data_as(C.POINTER(C.c_double)), n, stream))
        return re + 1j * im if complex_out else re


class Evaluator:
    """Phantom evaluator free functions (evaluate.cuh:32-304) as methods."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.L = _lib()

    def add_inplace(self, a, b, stream=None):
        check(self.L.shm_add_inplace(self.ctx.h, a.h, b.h, stream))

    def sub_inplace(self, a, b, stream=None):
        check(self.L.shm_sub_inplace(self.ctx.h, a.h, b.h, stream))

    def negate_inplace(self, a, stream=None):
        check(self.L.shm_negate_inplace(self.ctx.h, a.h, stream))

    def add_plain_inplace(self, a, pt, stream=None):
        check(self.L.shm_add_plain_inplace(self.ctx.h, a.h, pt.h, stream))

    def sub_plain_inplace(self, a, pt, stream=None):
        check(self.L.shm_sub_plain_inplace(self.ctx.h, a.h, pt.h, stream))

    def multiply_plain_inplace(self, ct, pt, stream=None):
        check(self.L.shm_multiply_plain_inplace(self.ctx.h, ct.h, pt.h, stream))

    def multiply_inplace(self, a, b, stream=None):
        check(self.L.shm_multiply_inplace(self.ctx.h, a.h, b.h, stream))

    def square_inplace(self, a, stream=None):
        check(self.L.shm_square_inplace(self.ctx.h, a.h, stream))

    def relinearize_inplace(self, ct, relin_key: KSwitchKey, stream=None):
        check(self.L.shm_relinearize_inplace(self.ctx.h, ct.h, relin_key.h, stream))

    def apply_kswitch_key_inplace(self, ct, key: KSwitchKey, stream=None):
        check(self.L.shm_apply_kswitch_key_inplace(self.ctx.h, ct.h, key.h, stream))

    def evaluate_chebyshev_inplace(self, cts, coeffs, lo: float, hi: float, relin_key: KSwitchKey, stream=None):
        """Polynomial::homomorphic_poly_evaluation (bootstrapping/common/Polynomial.cu:253-476)
        in Chebyshev form: every ciphertext of `cts` (one, or a list of up to 32 at
        one level and scale) becomes sum_k coeffs[k] T_k((2x - lo - hi) / (hi - lo))
        at its own scale, the list in lockstep."""
        cts = [cts] if isinstance(cts, Ciphertext) else list(cts)
        c = np.ascontiguousarray(coeffs, np.float64)
        hs = (C.c_void_p * len(cts))(*[x.h.value for x in cts])
        check(self.L.shm_evaluate_chebyshev(self.ctx.h, hs, len(cts), c.ctypes.data_as(C.POINTER(C.c_double)),
                                            c.size, float(lo), float(hi), relin_key.h, stream))

    def rescale_to_next_inplace(self, ct, stream=None):
        check(self.L.shm_rescale_to_next_inplace(self.ctx.h, ct.h, stream))

    def rescale_to_next_inplace_noalloc(self, ct, stream=None):
        """rescale_to_next_inplace without device allocation (scratch in a static device array)."""
        check(self.L.shm_rescale_to_next_inplace_noalloc(self.ctx.h, ct.h, stream))

    def mod_switch_to_next_inplace(self, ct, stream=None):
        check(self.L.shm_mod_switch_to_next_inplace(self.ctx.h, ct.h, stream))

    def mod_switch_to_inplace(self, ct, chain_index: int, stream=None):
        check(self.L.shm_mod_switch_to_inplace(self.ctx.h, ct.h, chain_index, stream))

    def plain_mod_switch_to_inplace(self, pt, chain_index: int, stream=None):
        check(self.L.shm_plain_mod_switch_to_inplace(self.ctx.h, pt.h, chain_index, stream))

    def apply_galois_inplace(self, ct, elt: int, gk: GaloisKeys, stream=None):
        check(self.L.shm_apply_galois_inplace(self.ctx.h, ct.h, elt, gk.h, stream))

    def rotate_vector_inplace(self, ct, step: int, gk: GaloisKeys, stream=None):
        check(self.L.shm_rotate_vector_inplace(self.ctx.h, ct.h, step, gk.h, stream))

    def complex_conjugate_inplace(self, ct, gk: GaloisKeys, stream=None):
        check(self.L.shm_complex_conjugate_inplace(self.ctx.h, ct.h, gk.h, stream))

    def hoisting_inplace(self, ct, gk: GaloisKeys, steps: Sequence[int], stream=None):
        """evaluate.cu:1727-1920: ct <- sum of rot(ct, step) over steps (one ModUp, one ModDown)."""
        arr = (C.c_int * len(steps))(*steps)
        check(self.L.shm_hoisting_inplace(self.ctx.h, ct.h, gk.h, arr, len(steps), stream))

    # moai::Evaluator facade operations (ckks_evaluator_parallel.cuh:444-561)
    def multiply_const_inplace(self, ct, value, stream=None):
        if isinstance(value, complex):
            check(self.L.shm_multiply_const_complex_inplace(self.ctx.h, ct.h, value.real, value.imag, stream))
        else:
            check(self.L.shm_multiply_const_inplace(self.ctx.h, ct.h, float(value), stream))

    def add_const_inplace(self, ct, value: float, stream=None):
        check(self.L.shm_add_const_inplace(self.ctx.h, ct.h, float(value), stream))

    def multiply_vector_inplace_reduced_error(self, ct, values, stream=None):
        v = np.asarray(values)
        re = np.ascontiguousarray(v.real, np.float64)
        im = np.ascontiguousarray(v.imag, np.float64) if np.iscomplexobj(v) else None
        check(self.L.shm_multiply_vector_inplace_reduced_error(
            self.ctx.h, ct.h, re.ctypes.data_as(C.POINTER(C.c_double)),
            im.ctypes.data_as(C.POINTER(C.c_double)) if im is not None else None, re.size, stream))

    def add_inplace_reduced_error(self, ct1, ct2, stream=None):
        check(self.L.shm_add_inplace_reduced_error(self.ctx.h, ct1.h, ct2.h, stream))

    def sub_inplace_reduced_error(self, ct1, ct2, stream=None):
        check(self.L.shm_sub_inplace_reduced_error(self.ctx.h, ct1.h, ct2.h, stream))

    def multiply_inplace_reduced_error(self, ct1, ct2, relin_key: KSwitchKey, stream=None):
        check(self.L.shm_multiply_inplace_reduced_error(self.ctx.h, ct1.h, ct2.h, relin_key.h, stream))

    def double_inplace(self, ct, stream=None):
        check(self.L.shm_double_inplace(self.ctx.h, ct.h, stream))

    def multiply_by_i_inplace(self, ct, negate: bool = False, stream=None):
        """Multiply every slot by i (X^(N/2)), or by -i (X^(3N/2)) when negate."""
        check(self.L.shm_multiply_by_i_inplace(self.ctx.h, ct.h, 1 if negate else 0, stream))


class Bootstrapper:
    """Slot-to-coefficient-first bootstrapping (Bootstrapper.cuh:63 +
    BatchBootstrap.cuh:674-709 bootstrap_ordinary_s2c_first).

    Inputs carry exactly 4 primes. rotation_steps() lists the Galois steps
    the transforms need (0 = conjugation) for create_galois_keys_from_steps.
    """

    def __init__(self, ctx: Context | None, layer3_factor: float = 1.0 / 16, K: int = 25, deg: int = 63,
                 double_angles: int = 2, input_bound: float = 1.0, s2c_input_primes: int = 4,
                 plan_only: tuple[int, int] | None = None):
        """plan_only=(log_n, size_q) builds the host plan without a context (no device)."""
        self.ctx = ctx
        p = capi.BootParams(s2c_input_primes, layer3_factor, K, deg, double_angles, input_bound)
        h = C.c_void_p()
        if plan_only is not None:
            check(_lib().shm_bootstrapper_create_plan(plan_only[0], plan_only[1], C.byref(p), C.byref(h)))
            self.n = 1 << (plan_only[0] - 1)
        else:
            check(_lib().shm_bootstrapper_create(ctx.h, C.byref(p), C.byref(h)))
            self.n = ctx.N // 2
        self.h = h
        self.K = K

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _lib().shm_bootstrapper_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def rotation_steps(self) -> list[int]:
        n = C.c_size_t()
        check(_lib().shm_bootstrapper_rotation_steps(self.h, None, 0, C.byref(n)))
        arr = (C.c_int * n.value)()
        check(_lib().shm_bootstrapper_rotation_steps(self.h, arr, n.value, C.byref(n)))
        return list(arr)

    def layer_info(self, transform: int, layer: int) -> dict:
        info = (C.c_int64 * 5)()
        check(_lib().shm_bootstrapper_layer_info(self.h, transform, layer, info))
        return {"base": info[0], "k0": info[1], "g": info[2], "giants": info[3], "chain": info[4]}

    def layer_diagonal(self, transform: int, layer: int, giant: int, baby: int) -> np.ndarray:
        n = self.n
        re = np.zeros(n)
        im = np.zeros(n)
        check(_lib().shm_bootstrapper_layer_diagonal(self.h, transform, layer, giant, baby,
                                                     re.ctypes.data_as(C.POINTER(C.c_double)),
                                                     im.ctypes.data_as(C.POINTER(C.c_double)), n))
        return re + 1j * im

    def evalmod_coeffs(self) -> np.ndarray:
        n = C.c_size_t()
        check(_lib().shm_bootstrapper_evalmod_coeffs(self.h, None, 0, C.byref(n)))
        out = np.zeros(n.value)
        check(_lib().shm_bootstrapper_evalmod_coeffs(self.h, out.ctypes.data_as(C.POINTER(C.c_double)), n.value,
                                                     C.byref(n)))
        return out

    def prepare(self, stream=None):
        check(_lib().shm_bootstrapper_prepare(self.h, stream))

    def slot_to_coeff(self, ct: Ciphertext, gk: GaloisKeys, stream=None):
        check(_lib().shm_bootstrap_slot_to_coeff(self.h, ct.h, gk.h, stream))

    def modraise_inplace(self, ct: Ciphertext, stream=None):
        check(_lib().shm_modraise_inplace(self.ctx.h, ct.h, stream))

    def coeff_to_slot(self, ct: Ciphertext, gk: GaloisKeys, stream=None):
        re, im = Ciphertext(self.ctx, 2, 1), Ciphertext(self.ctx, 2, 1)
        check(_lib().shm_bootstrap_coeff_to_slot(self.h, ct.h, re.h, im.h, gk.h, stream))
        return re, im

    def eval_mod(self, ct: Ciphertext, relin_key: KSwitchKey, stream=None):
        check(_lib().shm_bootstrap_eval_mod(self.h, ct.h, relin_key.h, stream))

    def eval_mod_batch(self, cts: Sequence[Ciphertext], relin_key: KSwitchKey, stream=None):
        """EvalMod of up to 32 ciphertexts at one level and scale, in lockstep."""
        hs = (C.c_void_p * len(cts))(*[c.h.value for c in cts])
        check(_lib().shm_bootstrap_eval_mod_batch(self.h, hs, len(cts), relin_key.h, stream))

    def shared_slot_to_coeff(self, keys: "SharedAKeys", batch: "SharedABatch", stream=None):
        check(_lib().shm_bootstrap_shared_slot_to_coeff(self.h, keys.h, batch.h, stream))

    def shared_coeff_to_slot(self, keys: "SharedAKeys", batch: "SharedABatch", stream=None):
        re, im = SharedABatch(self.ctx, keys.n, 1), SharedABatch(self.ctx, keys.n, 1)
        check(_lib().shm_bootstrap_shared_coeff_to_slot(self.h, keys.h, batch.h, re.h, im.h, stream))
        return re, im

    def bootstrap_s2c_first(self, cts: Sequence[Ciphertext], gk: GaloisKeys, relin_key: KSwitchKey,
                            out_scale: float = 0.0, stream=None, out=None) -> list:
        """Enqueue-only; `out` reuses existing output ciphertexts."""
        out = [Ciphertext(self.ctx, 2, 1) for _ in cts] if out is None else out
        ins = (C.c_void_p * len(cts))(*[c.h.value for c in cts])
        outs = (C.c_void_p * len(cts))(*[o.h.value for o in out])
        check(_lib().shm_bootstrap_s2c_first(self.h, ins, len(cts), outs, gk.h, relin_key.h, out_scale, stream))
        return out


class BootstrapServer:
    """Request server with continuous batching of S2C-first bootstraps (no counterpart
    in the reference). submit() queues a request of a few 4-prime ciphertexts and
    returns its id; wait(id) blocks until its outputs are ready and returns them with
    the request latency in ms. Requests from any thread share groups of up to
    max_group ciphertexts. Admission is bounded by the device's free memory and by
    mem_budget bytes (0: free memory only); `workers` (1..4) run groups concurrently.
    submit_packed() queues one input's column ciphertexts with data in slots
    [0, slots); requests of one column count and scale share ciphertexts, and wait()
    also returns the request's slot offset in the shared outputs."""

    def __init__(self, boot: Bootstrapper, gk: GaloisKeys, relin_key: KSwitchKey, mem_budget: int = 16 << 30,
                 max_group: int = 8, workers: int = 1):
        self.ctx = boot.ctx
        self._keep = (boot, gk, relin_key)
        h = C.c_void_p()
        check(_lib().shm_server_create_ex(boot.h, gk.h, relin_key.h, int(mem_budget), int(max_group), int(workers),
                                          C.byref(h)))
        self.h = h
        self._outs = {}

    def close(self):
        if getattr(self, "h", None):
            check(_lib().shm_server_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, cts: Sequence[Ciphertext]) -> int:
        outs = [Ciphertext(self.ctx, 2, 1) for _ in cts]
        rid = C.c_uint64()
        check(_lib().shm_server_submit(self.h, _ptrs(cts), len(cts), _ptrs(outs), C.byref(rid)))
        self._outs[rid.value] = (outs, False)
        return rid.value

    def submit_packed(self, cts: Sequence[Ciphertext], slots: int) -> int:
        outs = [Ciphertext(self.ctx, 2, 1) for _ in cts]
        rid = C.c_uint64()
        check(_lib().shm_server_submit_packed(self.h, _ptrs(cts), len(cts), int(slots), _ptrs(outs), C.byref(rid)))
        self._outs[rid.value] = (outs, True)
        return rid.value

    def wait(self, request_id: int):
        """Returns (outputs, latency_ms), or (outputs, latency_ms, slot_offset) for a packed request."""
        lat = C.c_double()
        outs, packed = self._outs.pop(request_id)
        if packed:
            off = C.c_uint32()
            check(_lib().shm_server_wait_packed(self.h, request_id, C.byref(lat), C.byref(off)))
            return outs, lat.value, off.value
        check(_lib().shm_server_wait(self.h, request_id, C.byref(lat)))
        return outs, lat.value

    def stats(self) -> dict:
        st = capi.ServerStats()
        check(_lib().shm_server_get_stats(self.h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in st._fields_}


def ct_pt_matrix_mul_wo_pre(enc_X: Sequence[Ciphertext], W: np.ndarray, col_X: int, col_W: int, row_W: int,
                            context: Context, variant: int = MATMUL_INT8, out=None, stream=None):
    """Ct_pt_matrix_mul.cuh:18-185: out[i] = rescale(sum_j X[j] (.) encode(W[j][i])).

    W is indexed W[j][i] (row_W x col_W). Returns col_W ciphertexts one level
    down with the scale of enc_X[0]. Raises ShmInvalidArgument on the
    reference's dimension errors.
    """
    if col_X != row_W:
        raise capi.ShmInvalidArgument(capi.SHM_ERR_INVALID_ARGUMENT, "bad dimensions of X or W")
    Wm = np.ascontiguousarray(np.asarray(W, np.float64)[:row_W, :col_W])
    if len(enc_X) < row_W:
        raise capi.ShmInvalidArgument(capi.SHM_ERR_INVALID_ARGUMENT, "enc_X size < row_W")
    chain = enc_X[0].chain_index
    if out is None:
        out = [Ciphertext(context, 2, chain + 1) for _ in range(col_W)]
    xs = (C.c_void_p * row_W)(*[x.h.value for x in enc_X[:row_W]])
    os_ = (C.c_void_p * col_W)(*[o.h.value for o in out])
    check(_lib().shm_ct_pt_matrix_mul(context.h, xs, row_W, Wm.ctypes.data_as(C.POINTER(C.c_double)), col_W, os_,
                                      variant, stream))
    return out


def ct_pt_matrix_mul_wo_pre_large(enc_X: Sequence[Ciphertext], W: np.ndarray, col_X: int, col_W: int, row_W: int,
                                  context: Context, variant: int = MATMUL_INT8, out=None, stream=None):
    """Ct_pt_matrix_mul.cuh:584-700, the wide-column form (col_W = 128 * col_W_t).

    The reference visits output c = i * col_W_t + k with weight column c, so the
    sums are those of ct_pt_matrix_mul_wo_pre. It fills only 128 * (col_W // 128)
    outputs; here a col_W that is not a multiple of 128 raises instead.
    """
    if col_W % 128:
        raise capi.ShmInvalidArgument(capi.SHM_ERR_INVALID_ARGUMENT, "col_W must be a multiple of 128")
    return ct_pt_matrix_mul_wo_pre(enc_X, W, col_X, col_W, row_W, context, variant, out, stream)


def ct_pt_matrix_mul_wo_pre_w_mask(enc_X: Sequence[Ciphertext], W: np.ndarray, bias_vec, col_X: int, col_W: int,
                                   row_W: int, context: Context, out=None, stream=None):
    """Ct_pt_matrix_mul.cuh:846-985 (and _w_mask_single, 987-1100): as
    ct_pt_matrix_mul_wo_pre, with the plaintext of term (j, i) encoding
    W[j][i] in the slots where bias_vec == 1 and 0 elsewhere (bias_vec holds
    slot_count = N/2 entries)."""
    if col_X != row_W:
        raise capi.ShmInvalidArgument(capi.SHM_ERR_INVALID_ARGUMENT, "bad dimensions of X or W")
    if len(enc_X) < row_W:
        raise capi.ShmInvalidArgument(capi.SHM_ERR_INVALID_ARGUMENT, "enc_X size < row_W")
    Wm = np.ascontiguousarray(np.asarray(W, np.float64)[:row_W, :col_W])
    bv = np.ascontiguousarray(bias_vec, np.int32)
    chain = enc_X[0].chain_index
    if out is None:
        out = [Ciphertext(context, 2, chain + 1) for _ in range(col_W)]
    xs = (C.c_void_p * row_W)(*[x.h.value for x in enc_X[:row_W]])
    os_ = (C.c_void_p * col_W)(*[o.h.value for o in out])
    check(_lib().shm_ct_pt_matrix_mul_masked(context.h, xs, row_W, Wm.ctypes.data_as(C.POINTER(C.c_double)), col_W,
                                             bv.ctypes.data_as(C.POINTER(C.c_int32)), bv.size, os_, stream))
    return out


def ct_ct_matrix_mul_colpacking(enc_X: Sequence[Ciphertext], enc_W: Sequence[Ciphertext], RotK: GaloisKeys,
                                relin_keys: KSwitchKey, context: Context, col_X: int, row_X: int, col_W: int,
                                row_W: int, num_batch: int, out=None, stream=None):
    """Ct_ct_matrix_mul.cuh:10-181: output[i] (i < row_X) =
    rescale(relinearize(sum_j enc_X[j] * rotate(enc_W[j], i * num_batch))) for the
    col_X column-packed ciphertexts of X and W. Raises ShmInvalidArgument when
    col_X != col_W or row_X != row_W (where the reference prints an error)."""
    if col_X != col_W or row_X != row_W:
        raise capi.ShmInvalidArgument(capi.SHM_ERR_INVALID_ARGUMENT, "bad dimensions of X or W")
    if len(enc_X) < col_X or len(enc_W) < col_X:
        raise capi.ShmInvalidArgument(capi.SHM_ERR_INVALID_ARGUMENT, "enc_X / enc_W hold fewer than col_X ciphertexts")
    chain = enc_X[0].chain_index
    if out is None:
        out = [Ciphertext(context, 2, chain + 1) for _ in range(row_X)]
    xs = (C.c_void_p * col_X)(*[x.h.value for x in enc_X[:col_X]])
    ws = (C.c_void_p * col_X)(*[w.h.value for w in enc_W[:col_X]])
    os_ = (C.c_void_p * row_X)(*[o.h.value for o in out])
    check(_lib().shm_ct_ct_matrix_mul_colpacking(context.h, xs, ws, col_X, row_X, col_W, row_W, num_batch, RotK.h,
                                                  relin_keys.h, os_, stream))
    return out


def ct_ct_matrix_mul_diagpacking(enc_X: Sequence[Ciphertext], enc_W: Sequence[Ciphertext], RotK: GaloisKeys,
                                 relin_keys: KSwitchKey, context: Context, col_X: int, row_X: int, col_W: int,
                                 row_W: int, num_batch: int, out=None, stream=None):
    """Ct_ct_matrix_mul.cuh:182-609 (baby-step giant-step): output[i] (i < col_W) =
    sum_j rotate(giant_j, j g num_batch), giant_j = rescale(relinearize(sum_k
    rotate(enc_W[i], k num_batch) * rotate(enc_X[j g + k], (col_X - j g) num_batch))),
    g = ceil(sqrt(col_X)); outputs carry enc_X's scale, as in the reference."""
    if len(enc_X) < col_X or len(enc_W) < col_W:
        raise capi.ShmInvalidArgument(capi.SHM_ERR_INVALID_ARGUMENT, "enc_X / enc_W hold too few ciphertexts")
    chain = enc_X[0].chain_index
    if out is None:
        out = [Ciphertext(context, 2, chain + 1) for _ in range(col_W)]
    xs = (C.c_void_p * col_X)(*[x.h.value for x in enc_X[:col_X]])
    ws = (C.c_void_p * col_W)(*[w.h.value for w in enc_W[:col_W]])
    os_ = (C.c_void_p * col_W)(*[o.h.value for o in out])
    check(_lib().shm_ct_ct_matrix_mul_diagpacking(context.h, xs, ws, col_X, row_X, col_W, row_W, num_batch, RotK.h,
                                                   relin_keys.h, os_, stream))
    return out


class Stream:
    """A created HIP stream (default priority, non-blocking); pass `.h` as `stream=`."""

    def __init__(self):
        h = C.c_void_p()
        check(_lib().shm_stream_create(C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            check(_lib().shm_stream_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """Calls recorded on a Stream between begin() and end(), replayed by launch()."""

    def __init__(self, stream: Stream):
        self.stream = stream
        self.h = None

    def begin(self):
        check(_lib().shm_graph_capture_begin(self.stream.h))

    def end(self):
        h = C.c_void_p()
        check(_lib().shm_graph_capture_end(self.stream.h, C.byref(h)))
        self.h = h

    def launch(self):
        check(_lib().shm_graph_launch(self.h, self.stream.h))

    def close(self):
        if getattr(self, "h", None):
            check(_lib().shm_graph_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def probe_fused_ntt_begin() -> None:
    """Arm the timing probe of the fused last NTT pass (ModDown combine, rescale)."""
    check(_lib().shm_probe_fused_ntt_begin())


def probe_fused_ntt_end() -> dict:
    """Launch count, summed kernel ms and algorithmic bytes of the probed launches."""
    n, ms, by = C.c_uint64(), C.c_double(), C.c_double()
    check(_lib().shm_probe_fused_ntt_end(C.byref(n), C.byref(ms), C.byref(by)))
    return {"launches": n.value, "ms": ms.value, "bytes": by.value}


def probe_plain_ntt_last() -> dict:
    """The plain forward NTT tile passes of the last probe window: launches, ms, bytes."""
    n, ms, by = C.c_uint64(), C.c_double(), C.c_double()
    check(_lib().shm_probe_plain_ntt_last(C.byref(n), C.byref(ms), C.byref(by)))
    return {"launches": n.value, "ms": ms.value, "bytes": by.value}


def keyswitch_counters(cap: int = 128) -> dict:
    """Key-switching work per level since the last reset: {Ql: (modup, inner, moddown)}."""
    a, b, c = (np.zeros(cap, np.uint64) for _ in range(3))
    check(_lib().shm_keyswitch_counters(_u64p(a), _u64p(b), _u64p(c), cap))
    return {q: (int(a[q]), int(b[q]), int(c[q])) for q in range(cap) if a[q] or b[q] or c[q]}


def keyswitch_counters_reset() -> None:
    check(_lib().shm_keyswitch_counters_reset())


def set_tuning(name: str, value: int) -> None:
    """Set a process-wide knob; results never depend on them.

    "ks_batch_slots_log2", "ks_multi_slots_log2": slots of the key inner products (0..4, -1 = autotuned);
    "matmul_overlap": 0/1; "ntt_tile_threads": 512 (radix-8 rounds) or 256 (radix-16 rounds) threads per
    NTT tile block."""
    check(_lib().shm_set_tuning(name.encode(), int(value)))


def autotune_enable(on: bool = True) -> None:
    """Switch the inner-product autotuner on or off (on by default)."""
    check(_lib().shm_autotune_enable(1 if on else 0))


def autotune_clear() -> None:
    check(_lib().shm_autotune_clear())


def autotune_report() -> str:
    """The autotuned shapes, the chosen slot count and the candidates' timings."""
    need = C.c_size_t()
    check(_lib().shm_autotune_report(None, 0, C.byref(need)))
    buf = C.create_string_buffer(need.value)
    check(_lib().shm_autotune_report(buf, need.value, C.byref(need)))
    return buf.value.decode()


def get_tuning(name: str) -> int:
    v = C.c_int()
    check(_lib().shm_get_tuning(name.encode(), C.byref(v)))
    return v.value


def synchronize(stream=None) -> None:
    """Wait for all work queued on `stream` (None: the default stream)."""
    check(_lib().shm_synchronize(stream))


def last_matmul_kernel_ms() -> float:
    return float(_lib().shm_last_matmul_kernel_ms())


def last_matmul_rescale_tile_ms() -> float:
    return float(_lib().shm_last_matmul_rescale_tile_ms())


def last_matmul_variant() -> int:
    return int(_lib().shm_last_matmul_variant())


def _ptrs(objs):
    return (C.c_void_p * len(objs))(*[o.h.value for o in objs])


class SharedABatch:
    """SharedACiphertextBatch (SharedACiphertext.cuh:22-99): n members (b_i, a)."""

    def __init__(self, ctx: Context, n: int, chain_index: int = 1):
        self.ctx, self.n = ctx, n
        h = C.c_void_p()
        check(_lib().shm_shared_a_batch_create(ctx.h, n, chain_index, C.byref(h)))
        self.h = h

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _lib().shm_shared_a_batch_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def info(self):
        n, ch, cm, sc = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_double()
        check(_lib().shm_shared_a_batch_info(self.h, C.byref(n), C.byref(ch), C.byref(cm), C.byref(sc)))
        return n.value, ch.value, cm.value, sc.value

    @property
    def chain_index(self):
        return self.info()[1]

    @property
    def coeff_modulus_size(self):
        return self.info()[2]

    @property
    def scale(self):
        return self.info()[3]

    @scale.setter
    def scale(self, v: float):
        check(_lib().shm_shared_a_batch_set_scale(self.h, v))

    def data_ptr(self) -> int:
        """Device address of the words [n+1][Ql][N] (b_0..b_{n-1}, then a)."""
        return _lib().shm_shared_a_batch_data(self.h)

    def words(self) -> int:
        return (self.n + 1) * self.coeff_modulus_size * self.ctx.N

    def load(self, limbs: np.ndarray, stream=None) -> "SharedABatch":
        a = np.ascontiguousarray(limbs, np.uint64)
        check(_lib().shm_shared_a_batch_import(self.h, _u64p(a), a.size, stream))
        return self

    def numpy(self, stream=None) -> np.ndarray:
        out = np.zeros(self.words(), np.uint64)
        check(_lib().shm_shared_a_batch_export(self.h, _u64p(out), out.size, stream))
        return out

    def member(self, i: int, stream=None) -> Ciphertext:
        ct = Ciphertext(self.ctx, 2, 1)
        check(_lib().shm_shared_a_batch_member(self.h, i, ct.h, stream))
        return ct

    def add_inplace(self, other: "SharedABatch", stream=None):
        check(_lib().shm_shared_a_add_inplace(self.h, other.h, stream))

    def multiply_plain_inplace(self, pt: Plaintext, stream=None):
        check(_lib().shm_shared_a_multiply_plain_inplace(self.h, pt.h, stream))

    def rescale_to_next_inplace(self, stream=None):
        check(_lib().shm_shared_a_rescale_to_next_inplace(self.h, stream))

    def modraise_inplace(self, stream=None):
        check(_lib().shm_shared_a_modraise_inplace(self.h, stream))


class SharedAKeys:
    """SharedAKeyGen + SharedAFormatKey (FmtSwitch.cuh:28-105) and shared-a Galois
    keys (ShRotate.cuh:25-35), all resident in HBM."""

    def __init__(self, ctx: Context, sk: SecretKey, n: int, hamming_weight: int, seed: int, chain_index: int):
        self.ctx, self.n, self.chain_index = ctx, n, chain_index
        h = C.c_void_p()
        check(_lib().shm_shared_a_keys_create(ctx.h, sk.h, n, hamming_weight, seed, chain_index, C.byref(h)))
        self.h = h

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _lib().shm_shared_a_keys_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def add_galois(self, steps: Sequence[int], seed: int):
        arr = (C.c_int * len(steps))(*steps)
        check(_lib().shm_shared_a_keys_add_galois(self.h, arr, len(steps), seed))

    def aux_coeffs(self, i: int) -> np.ndarray:
        out = np.zeros(self.ctx.N, np.int8)
        check(_lib().shm_shared_a_keys_aux_coeffs(self.h, i, out.ctypes.data_as(C.POINTER(C.c_int8)), out.size))
        return out

    def aux_secret(self, i: int) -> SecretKey:
        return SecretKey(self.ctx, coeffs=self.aux_coeffs(i))

    def export_format(self):
        c = self.ctx
        qlp = c.ql(self.chain_index) + c.size_p
        a = np.zeros(self.n * qlp * c.N, np.uint64)
        B = np.zeros(self.n * self.n * qlp * c.N, np.uint64)
        check(_lib().shm_shared_a_keys_export_format(self.h, _u64p(a), _u64p(B), B.size))
        return a, B

    def export_galois(self, elt: int):
        c = self.ctx
        A = np.zeros(c.dnum * (c.size_q + c.size_p) * c.N, np.uint64)
        B = np.zeros(c.dnum * self.n * (c.size_q + c.size_p) * c.N, np.uint64)
        check(_lib().shm_shared_a_keys_export_galois(self.h, elt, _u64p(A), _u64p(B), B.size))
        return A, B

    def fmt_switch_forward(self, cts: Sequence[Ciphertext], stream=None) -> SharedABatch:
        out = SharedABatch(self.ctx, self.n, self.chain_index)
        check(_lib().shm_fmt_switch_forward(self.h, _ptrs(cts), len(cts), out.h, stream))
        return out

    def fmt_switch_backward(self, batch: SharedABatch, stream=None) -> list:
        out = [Ciphertext(self.ctx, 2, 1) for _ in range(batch.n)]
        check(_lib().shm_fmt_switch_backward(self.h, batch.h, _ptrs(out), len(out), stream))
        return out

    def rotate_inplace(self, batch: SharedABatch, step: int, stream=None):
        check(_lib().shm_shared_a_rotate_inplace(self.h, batch.h, step, stream))


def bootstrap_shared_a_s2c_first(boot: "Bootstrapper", keys: SharedAKeys, cts: Sequence[Ciphertext],
                                 relin_key: KSwitchKey, out_scale: float = 0.0, stream=None, out=None) -> list:
    """BatchBootstrap.cuh:719-820: shared-a S2C-first bootstrap of exactly keys.n ciphertexts
    (enqueue-only; `out` reuses existing output ciphertexts)."""
    out = [Ciphertext(keys.ctx, 2, 1) for _ in cts] if out is None else out
    check(_lib().shm_bootstrap_shared_a_s2c_first(boot.h, keys.h, _ptrs(cts), len(cts), _ptrs(out), relin_key.h,
                                                  out_scale, stream))
    return out


def bootstrap_shared_a_s2c_first_chunks(boot: "Bootstrapper", keys: SharedAKeys, cts: Sequence[Ciphertext],
                                        relin_key: KSwitchKey, out_scale: float = 0.0, stream=None,
                                        out=None) -> list:
    """The shared-a bootstrap over k * keys.n ciphertexts, k chunks in order (enqueue-only)."""
    out = [Ciphertext(keys.ctx, 2, 1) for _ in cts] if out is None else out
    check(_lib().shm_bootstrap_shared_a_s2c_first_chunks(boot.h, keys.h, _ptrs(cts), len(cts), _ptrs(out),
                                                         relin_key.h, out_scale, stream))
    return out


class BootGraph:
    """One bootstrap chunk recorded into a HIP graph on a created stream and replayed
    per chunk of a longer input list (shm_boot_graph_*). Build with shared_a() or
    ordinary(); run() only enqueues (ordered against the default stream)."""

    def __init__(self, h, ctx, keepalive):
        self.h, self.ctx, self._keep = h, ctx, keepalive
        c, nodes = C.c_uint32(), C.c_uint64()
        check(_lib().shm_boot_graph_info(self.h, C.byref(c), C.byref(nodes)))
        self.chunk, self.nodes = c.value, nodes.value

    @classmethod
    def shared_a(cls, boot: "Bootstrapper", keys: SharedAKeys, relin_key: KSwitchKey, in_scale: float,
                 out_scale: float = 0.0, stream=None):
        stream = stream or Stream()
        h = C.c_void_p()
        check(_lib().shm_boot_graph_create_shared_a(boot.h, keys.h, relin_key.h, float(in_scale), float(out_scale),
                                                    stream.h, C.byref(h)))
        return cls(h, boot.ctx, (boot, keys, relin_key, stream))

    @classmethod
    def ordinary(cls, boot: "Bootstrapper", gk: GaloisKeys, relin_key: KSwitchKey, group: int, in_scale: float,
                 out_scale: float = 0.0, stream=None):
        stream = stream or Stream()
        h = C.c_void_p()
        check(_lib().shm_boot_graph_create_ordinary(boot.h, gk.h, relin_key.h, group, float(in_scale),
                                                    float(out_scale), stream.h, C.byref(h)))
        return cls(h, boot.ctx, (boot, gk, relin_key, stream))

    def run(self, cts: Sequence[Ciphertext], out=None) -> list:
        out = [Ciphertext(self.ctx, 2, 1) for _ in cts] if out is None else out
        check(_lib().shm_boot_graph_run(self.h, _ptrs(cts), len(cts), _ptrs(out)))
        return out

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _lib().shm_boot_graph_destroy(self.h)
                self.h = None
        except Exception:
            pass
