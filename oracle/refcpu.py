"""ctypes front end of the oracle (oracle/librefcpu.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the shmoai_amd product package.
Functions mirror refcpu.h; arrays are numpy uint64/int64/float64, C-contiguous.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "librefcpu.so")
_lib = None

u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    """Compile oracle/librefcpu.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp = C.c_void_p
        sz = C.c_size_t
        sig = {
            "ref_ctx_new": (vp, [C.c_int, C.c_int, C.POINTER(C.c_int), C.c_int]),
            "ref_ctx_free": (None, [vp]),
            "ref_ctx_nq": (sz, [vp]),
            "ref_ctx_np": (sz, [vp]),
            "ref_ctx_primes": (None, [vp, u64p]),
            "ref_ctx_psi": (None, [vp, u64p]),
            "ref_ctx_tables": (None, [vp, sz, u64p, u64p]),
            "ref_set_threads": (None, [C.c_int]),
            "ref_is_prime": (C.c_int, [C.c_uint64]),
            "ref_get_primes": (None, [sz, C.c_int, sz, u64p]),
            "ref_ntt": (None, [vp, u64p, sz, u32p]),
            "ref_intt": (None, [vp, u64p, sz, u32p]),
            "ref_add": (None, [vp, u64p, u64p, u64p, sz, sz]),
            "ref_sub": (None, [vp, u64p, u64p, u64p, sz, sz]),
            "ref_negate": (None, [vp, u64p, u64p, sz, sz]),
            "ref_multiply_plain": (None, [vp, u64p, u64p, sz, sz]),
            "ref_tensor": (None, [vp, u64p, u64p, u64p, sz]),
            "ref_rescale": (None, [vp, u64p, u64p, sz, sz]),
            "ref_mod_switch_drop": (None, [vp, u64p, u64p, sz, sz]),
            "ref_galois_elt_from_step": (C.c_uint32, [vp, C.c_int]),
            "ref_galois_table": (None, [vp, C.c_uint32, u32p]),
            "ref_apply_galois_ntt": (None, [vp, C.c_uint32, u64p, u64p, sz]),
            "ref_beta": (sz, [vp, sz]),
            "ref_modup": (None, [vp, u64p, sz, u64p]),
            "ref_ks_inner": (None, [vp, u64p, u64p, sz, u64p]),
            "ref_moddown": (None, [vp, u64p, sz, u64p]),
            "ref_keyswitch": (None, [vp, u64p, u64p, sz, u64p, u64p]),
            "ref_rotate": (None, [vp, u64p, sz, C.c_uint32, u64p, u64p]),
            "ref_relinearize": (None, [vp, u64p, sz, u64p, u64p]),
            "ref_splitmix64": (C.c_uint64, [C.c_uint64, C.c_uint64]),
            "ref_sample_uniform": (None, [vp, C.c_uint64, sz, u32p, u64p]),
            "ref_sample_gaussian": (None, [vp, C.c_uint64, i64p]),
            "ref_sample_ternary": (None, [vp, C.c_uint64, sz, i64p]),
            "ref_sk_ntt": (None, [vp, i64p, u64p]),
            "ref_gen_kswitch_key": (None, [vp, u64p, u64p, C.c_uint64, u64p]),
            "ref_gen_relin_key": (None, [vp, u64p, C.c_uint64, u64p]),
            "ref_gen_galois_key": (None, [vp, u64p, C.c_uint32, C.c_uint64, u64p]),
            "ref_encrypt_symmetric": (None, [vp, u64p, u64p, sz, C.c_uint64, u64p]),
            "ref_decrypt": (None, [vp, u64p, u64p, sz, sz, u64p]),
            "ref_sample_ternary_uniform": (None, [vp, C.c_uint64, i64p]),
            "ref_gen_public_key": (None, [vp, u64p, C.c_uint64, u64p]),
            "ref_encrypt_asymmetric": (None, [vp, u64p, u64p, C.c_uint64, u64p]),
            "ref_encode": (None, [vp, f64p, C.c_void_p, sz, sz, sz, C.c_double, u64p]),
            "ref_encode_scalar": (None, [vp, C.c_double, sz, C.c_double, u64p]),
            "ref_decode": (None, [vp, u64p, sz, sz, C.c_double, f64p, f64p]),
            "ref_ctpt_matmul": (None, [vp, C.POINTER(C.c_void_p), sz, f64p, sz, sz, f64p, sz, sz,
                                       C.POINTER(C.c_void_p)]),
            "ref_modraise": (None, [vp, u64p, u64p]),
            "ref_mul_monomial": (None, [vp, u64p, sz, sz, sz]),
            "ref_extend_exact": (None, [vp, u64p, sz, u64p]),
            "ref_gen_fmt_key": (None, [vp, u64p, C.POINTER(C.c_void_p), sz, sz, C.c_uint64, u64p, u64p]),
            "ref_fmt_switch_forward": (None, [vp, C.POINTER(C.c_void_p), sz, sz, u64p, u64p, u64p]),
            "ref_gen_shared_galois_key": (None, [vp, C.POINTER(C.c_void_p), sz, C.c_uint32, C.c_uint64, u64p,
                                                 u64p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    lib().ref_set_threads(int(n))


class Ref:
    """One CKKS parameter set of the oracle (primes, NTT tables, encoder roots)."""

    def __init__(self, logN: int, bit_sizes, n_special: int):
        self.L = lib()
        arr = (C.c_int * len(bit_sizes))(*bit_sizes)
        self.h = self.L.ref_ctx_new(logN, len(bit_sizes), arr, n_special)
        self.logN = logN
        self.N = 1 << logN
        self.nQ = int(self.L.ref_ctx_nq(self.h))
        self.nP = int(self.L.ref_ctx_np(self.h))
        self.primes = np.zeros(self.nQ + self.nP, np.uint64)
        self.L.ref_ctx_primes(self.h, self.primes)

    def __del__(self):
        try:
            if self.h:
                self.L.ref_ctx_free(self.h)
                self.h = None
        except Exception:
            pass

    # helpers ----------------------------------------------------------------
    @property
    def nQP(self):
        return self.nQ + self.nP

    def dnum(self):
        return self.nQ // self.nP

    def beta(self, Ql):
        return int(self.L.ref_beta(self.h, Ql))

    def pidx(self, Ql, withP=False):
        v = list(range(Ql)) + (list(range(self.nQ, self.nQ + self.nP)) if withP else [])
        return np.array(v, np.uint32)

    def psi(self):
        out = np.zeros(self.nQP, np.uint64)
        self.L.ref_ctx_psi(self.h, out)
        return out

    def tables(self, prime):
        tw = np.zeros(self.N, np.uint64)
        itw = np.zeros(self.N, np.uint64)
        self.L.ref_ctx_tables(self.h, prime, tw, itw)
        return tw, itw

    # transforms ---------------------------------------------------------------
    def ntt(self, data, pidx):
        d = np.ascontiguousarray(data, np.uint64).copy()
        self.L.ref_ntt(self.h, d, len(pidx), np.ascontiguousarray(pidx, np.uint32))
        return d

    def intt(self, data, pidx):
        d = np.ascontiguousarray(data, np.uint64).copy()
        self.L.ref_intt(self.h, d, len(pidx), np.ascontiguousarray(pidx, np.uint32))
        return d

    # elementwise ----------------------------------------------------------------
    def add(self, a, b, Ql, npolys):
        out = np.empty_like(a)
        self.L.ref_add(self.h, a, b, out, Ql, npolys)
        return out

    def sub(self, a, b, Ql, npolys):
        out = np.empty_like(a)
        self.L.ref_sub(self.h, a, b, out, Ql, npolys)
        return out

    def negate(self, a, Ql, npolys):
        out = np.empty_like(a)
        self.L.ref_negate(self.h, a, out, Ql, npolys)
        return out

    def multiply_plain(self, ct, pt, Ql, npolys=2):
        out = np.ascontiguousarray(ct, np.uint64).copy()
        self.L.ref_multiply_plain(self.h, out, np.ascontiguousarray(pt, np.uint64), Ql, npolys)
        return out

    def tensor(self, a, b, Ql):
        out = np.zeros(3 * Ql * self.N, np.uint64)
        self.L.ref_tensor(self.h, a, b, out, Ql)
        return out

    def rescale(self, ct, Ql, npolys=2):
        out = np.zeros(npolys * (Ql - 1) * self.N, np.uint64)
        self.L.ref_rescale(self.h, np.ascontiguousarray(ct, np.uint64), out, Ql, npolys)
        return out

    def mod_switch_drop(self, ct, Ql, npolys=2):
        out = np.zeros(npolys * (Ql - 1) * self.N, np.uint64)
        self.L.ref_mod_switch_drop(self.h, np.ascontiguousarray(ct, np.uint64), out, Ql, npolys)
        return out

    # galois ----------------------------------------------------------------------
    def elt(self, step):
        return int(self.L.ref_galois_elt_from_step(self.h, step))

    def galois_table(self, elt):
        t = np.zeros(self.N, np.uint32)
        self.L.ref_galois_table(self.h, elt, t)
        return t

    def apply_galois(self, data, elt, nlimbs):
        out = np.zeros(nlimbs * self.N, np.uint64)
        self.L.ref_apply_galois_ntt(self.h, elt, np.ascontiguousarray(data, np.uint64), out, nlimbs)
        return out

    # key switching -----------------------------------------------------------------
    def modup(self, c2, Ql):
        out = np.zeros(self.beta(Ql) * (Ql + self.nP) * self.N, np.uint64)
        self.L.ref_modup(self.h, np.ascontiguousarray(c2, np.uint64), Ql, out)
        return out

    def ks_inner(self, modup, key, Ql):
        if self.beta(Ql) > self.dnum():
            raise ValueError("key switching key has too few digits for this level")
        out = np.zeros(2 * (Ql + self.nP) * self.N, np.uint64)
        self.L.ref_ks_inner(self.h, modup, key, Ql, out)
        return out

    def moddown(self, cx, Ql):
        cx = np.ascontiguousarray(cx, np.uint64).copy()
        out = np.zeros(Ql * self.N, np.uint64)
        self.L.ref_moddown(self.h, cx, Ql, out)
        return out

    def keyswitch(self, c2, key, Ql):
        k0 = np.zeros(Ql * self.N, np.uint64)
        k1 = np.zeros(Ql * self.N, np.uint64)
        self.L.ref_keyswitch(self.h, np.ascontiguousarray(c2, np.uint64), key, Ql, k0, k1)
        return k0, k1

    def rotate(self, ct, Ql, elt, key):
        out = np.zeros(2 * Ql * self.N, np.uint64)
        self.L.ref_rotate(self.h, np.ascontiguousarray(ct, np.uint64), Ql, elt, key, out)
        return out

    def relinearize(self, ct3, Ql, key):
        out = np.zeros(2 * Ql * self.N, np.uint64)
        self.L.ref_relinearize(self.h, np.ascontiguousarray(ct3, np.uint64), Ql, key, out)
        return out

    # samplers / keys / encryption -----------------------------------------------------
    def splitmix64(self, key, ctr):
        return int(self.L.ref_splitmix64(key, ctr))

    def sample_uniform(self, seed, pidx):
        out = np.zeros(len(pidx) * self.N, np.uint64)
        self.L.ref_sample_uniform(self.h, seed, len(pidx), np.ascontiguousarray(pidx, np.uint32), out)
        return out

    def sample_gaussian(self, seed):
        out = np.zeros(self.N, np.int64)
        self.L.ref_sample_gaussian(self.h, seed, out)
        return out

    def sample_ternary(self, seed, hw):
        out = np.zeros(self.N, np.int64)
        self.L.ref_sample_ternary(self.h, seed, hw, out)
        return out

    def sk_ntt(self, coeffs):
        out = np.zeros(self.nQP * self.N, np.uint64)
        self.L.ref_sk_ntt(self.h, np.ascontiguousarray(coeffs, np.int64), out)
        return out

    def key_size(self):
        return self.dnum() * 2 * self.nQP * self.N

    def gen_kswitch_key(self, sk_qp, newkey_qp, seed):
        out = np.zeros(self.key_size(), np.uint64)
        self.L.ref_gen_kswitch_key(self.h, sk_qp, np.ascontiguousarray(newkey_qp, np.uint64), seed, out)
        return out

    def gen_relin_key(self, sk_qp, seed):
        out = np.zeros(self.key_size(), np.uint64)
        self.L.ref_gen_relin_key(self.h, sk_qp, seed, out)
        return out

    def gen_galois_key(self, sk_qp, elt, seed):
        out = np.zeros(self.key_size(), np.uint64)
        self.L.ref_gen_galois_key(self.h, sk_qp, elt, seed, out)
        return out

    def encrypt_symmetric(self, sk_qp, pt, Ql, seed):
        out = np.zeros(2 * Ql * self.N, np.uint64)
        self.L.ref_encrypt_symmetric(self.h, sk_qp, np.ascontiguousarray(pt, np.uint64), Ql, seed, out)
        return out

    def decrypt(self, sk_qp, ct, size, Ql):
        out = np.zeros(Ql * self.N, np.uint64)
        self.L.ref_decrypt(self.h, sk_qp, np.ascontiguousarray(ct, np.uint64), size, Ql, out)
        return out

    def sample_ternary_uniform(self, seed):
        out = np.zeros(self.N, np.int64)
        self.L.ref_sample_ternary_uniform(self.h, seed, out)
        return out

    def gen_public_key(self, sk_qp, seed):
        """secretkey.cu:524-535 gen_publickey: [2][QP][N] NTT form."""
        out = np.zeros(2 * self.nQP * self.N, np.uint64)
        self.L.ref_gen_public_key(self.h, np.ascontiguousarray(sk_qp, np.uint64), seed, out)
        return out

    def encrypt_asymmetric(self, pk, pt, seed):
        """secretkey.cu:145-217 encrypt_asymmetric of a plaintext at chain index 1."""
        out = np.zeros(2 * self.nQ * self.N, np.uint64)
        self.L.ref_encrypt_asymmetric(self.h, np.ascontiguousarray(pk, np.uint64),
                                      np.ascontiguousarray(pt, np.uint64), seed, out)
        return out

    # encoder -------------------------------------------------------------------------
    def encode(self, re, Ql, scale, im=None, sparse_slots=None):
        re = np.ascontiguousarray(re, np.float64)
        n = sparse_slots or self.N // 2
        out = np.zeros(Ql * self.N, np.uint64)
        imp = None
        if im is not None:
            im = np.ascontiguousarray(im, np.float64)
            imp = im.ctypes.data_as(C.c_void_p)
        self.L.ref_encode(self.h, re, imp, len(re), n, Ql, float(scale), out)
        return out

    def encode_scalar(self, value, Ql, scale):
        out = np.zeros(Ql * self.N, np.uint64)
        self.L.ref_encode_scalar(self.h, float(value), Ql, float(scale), out)
        return out

    def decode(self, pt, Ql, scale, sparse_slots=None):
        n = sparse_slots or self.N // 2
        re = np.zeros(n, np.float64)
        im = np.zeros(n, np.float64)
        self.L.ref_decode(self.h, np.ascontiguousarray(pt, np.uint64), Ql, n, float(scale), re, im)
        return re, im

    # MOAI ct-pt matmul ------------------------------------------------------------
    def ctpt_matmul(self, X, W, Ql, scales, col_begin=0, col_end=None):
        row_w, col_w = W.shape
        col_end = col_w if col_end is None else col_end
        Xs = [np.ascontiguousarray(x, np.uint64) for x in X]
        xp = (C.c_void_p * row_w)(*[x.ctypes.data for x in Xs])
        outs = [np.zeros(2 * (Ql - 1) * self.N, np.uint64) for _ in range(col_begin, col_end)]
        op = (C.c_void_p * len(outs))(*[o.ctypes.data for o in outs])
        self.L.ref_ctpt_matmul(self.h, xp, row_w, np.ascontiguousarray(W, np.float64), col_w, Ql,
                               np.ascontiguousarray(scales, np.float64), col_begin, col_end, op)
        return outs

    def ctpt_matmul_masked(self, X, W, mask, Ql, scales):
        """ct_pt_matrix_mul_wo_pre_w_mask (Ct_pt_matrix_mul.cuh:846-985): for each output
        i, sum over j of X[j] (.) encode({mask[t] == 1 ? W[j][i] : 0}, scale_j), each
        product and sum reduced as multiply_plain and add_inplace do, then rescale."""
        mask = np.asarray(mask)
        outs = []
        for i in range(W.shape[1]):
            acc = None
            for j in range(W.shape[0]):
                pt = self.encode(np.where(mask == 1, W[j, i], 0.0), Ql, scales[j])
                t = self.multiply_plain(X[j], pt, Ql)
                acc = t if acc is None else self.add(acc, t, Ql, 2)
            outs.append(self.rescale(acc, Ql))
        return outs

    # bootstrapping pieces -------------------------------------------------------------
    def ctct_matmul_colpacking(self, X, W, Ql, row_x, num_batch, gal_keys, relin_key):
        """Ct_ct_matrix_mul.cuh:10-181 ct_ct_matrix_mul_colpacking: output i is
        rescale(relinearize(sum_j X[j] * rotate(W[j], i * num_batch))), the products
        formed by multiply (tensor, no relinearisation, Ct_ct_matrix_mul.cuh:125-143)
        and summed by add_inplace before one relinearize and rescale (147-148).
        gal_keys maps each rotation step (mod the slot count) to its Galois key; a
        step of 0 is the identity."""
        n = self.N // 2
        out = []
        for i in range(row_x):
            step = (i * num_batch) % n
            acc = None
            for j in range(len(X)):
                w = W[j] if step == 0 else self.rotate(W[j], Ql, self.elt(step), gal_keys[step])
                t = self.tensor(np.ascontiguousarray(X[j], np.uint64), np.ascontiguousarray(w, np.uint64), Ql)
                acc = t if acc is None else self.add(acc, t, Ql, 3)
            out.append(self.rescale(self.relinearize(acc, Ql, relin_key), Ql))
        return out

    def ctct_matmul_diagpacking(self, X, W, Ql, num_batch, gal_keys, relin_key):
        """Ct_ct_matrix_mul.cuh:182-609 ct_ct_matrix_mul_diagpacking: g = ceil(sqrt(col_X)),
        b = ceil(col_X / g); X[j g + k] pre-rotated by (col_X - j g) num_batch for j > 0
        (241-258); babies rot(W[i], k num_batch) (342-349); giant j = rescale(relinearize(
        sum_k baby_k * X'[j g + k])) (351-372); output i = sum_j rot(giant_j, j g num_batch)
        (374-385). gal_keys maps each step (mod the slot count) to its key; 0 is the
        identity. Returns the output limbs (the scale label is the caller's)."""
        n, col_x = self.N // 2, len(X)
        g = int(np.sqrt(col_x))
        if g * g < col_x:
            g += 1
        b = col_x // g
        if b * g < col_x:
            b += 1

        def rot(ct, step, ql=Ql):
            step %= n
            return ct if step == 0 else self.rotate(ct, ql, self.elt(step), gal_keys[step])

        xr = [X[i] if i // g == 0 else rot(X[i], (col_x - (i // g) * g) * num_batch) for i in range(col_x)]
        out = []
        for w in W:
            baby = [w] + [rot(w, k * num_batch) for k in range(1, g)]
            acc_out = None
            for j in range(b):
                acc = None
                for k in range(g):
                    if j * g + k >= col_x:
                        break
                    t = self.tensor(np.ascontiguousarray(baby[k], np.uint64),
                                    np.ascontiguousarray(xr[j * g + k], np.uint64), Ql)
                    acc = t if acc is None else self.add(acc, t, Ql, 3)
                giant = self.rescale(self.relinearize(acc, Ql, relin_key), Ql)
                if j == 0:
                    acc_out = giant
                else:
                    acc_out = self.add(acc_out, rot(giant, j * g * num_batch, Ql - 1), Ql - 1, 2)
            out.append(acc_out)
        return out

    def modraise(self, ct_low):
        """[2][N] at prime 0 (NTT form) -> [2][nQ][N] (Bootstrapper.cu:3287-3356)."""
        out = np.zeros(2 * self.nQ * self.N, np.uint64)
        self.L.ref_modraise(self.h, np.ascontiguousarray(ct_low, np.uint64), out)
        return out

    def mul_monomial(self, data, nlimbs, npolys, power):
        d = np.ascontiguousarray(data, np.uint64).copy()
        self.L.ref_mul_monomial(self.h, d, nlimbs, npolys, power)
        return d

    # shared-a format ------------------------------------------------------------------
    @staticmethod
    def _pp(arrs):
        return (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])

    def extend_exact(self, coef, Ql):
        """FmtSwitch.cu:82-108: coefficient residues over Ql primes -> residues mod P primes."""
        out = np.zeros(self.nP * self.N, np.uint64)
        self.L.ref_extend_exact(self.h, np.ascontiguousarray(coef, np.uint64), Ql, out)
        return out

    def gen_fmt_key(self, sk_qp, aux_qp, Ql, seed):
        n = len(aux_qp)
        QlP = Ql + self.nP
        aux = [np.ascontiguousarray(a, np.uint64) for a in aux_qp]
        a_fmt = np.zeros(n * QlP * self.N, np.uint64)
        B = np.zeros(n * n * QlP * self.N, np.uint64)
        self.L.ref_gen_fmt_key(self.h, np.ascontiguousarray(sk_qp, np.uint64), self._pp(aux), n, Ql, seed, a_fmt, B)
        return a_fmt, B

    def fmt_switch_forward(self, cts, Ql, a_fmt, B):
        n = len(cts)
        cs = [np.ascontiguousarray(c, np.uint64) for c in cts]
        out = np.zeros((n + 1) * Ql * self.N, np.uint64)
        self.L.ref_fmt_switch_forward(self.h, self._pp(cs), n, Ql, np.ascontiguousarray(a_fmt, np.uint64),
                                      np.ascontiguousarray(B, np.uint64), out)
        return out

    def gen_shared_galois_key(self, aux_qp, elt, seed):
        n = len(aux_qp)
        aux = [np.ascontiguousarray(a, np.uint64) for a in aux_qp]
        dnum = self.nQ // self.nP
        A = np.zeros(dnum * self.nQP * self.N, np.uint64)
        B = np.zeros(dnum * n * self.nQP * self.N, np.uint64)
        self.L.ref_gen_shared_galois_key(self.h, self._pp(aux), n, elt, seed, A, B)
        return A, B
