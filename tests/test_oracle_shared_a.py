"""Oracle pins for the shared-a format (CPU only).

- exact Ql -> P extension (FmtSwitch.cu:82-108) equals x mod p computed with
  Python big integers from the CRT composition, for random and edge values;
  the reference's floor guard (1e-9) makes x > Q (1 - 1e-9) come out as x - Q,
  the centred representative, which the pin states explicitly;
- format key entries satisfy B[j][i] + a_fmt[j] s'_i - [i == j] P s = -e with a
  small e, and shared Galois key entries B_{g,i} + A_g s'_i - P [sigma(s'_i)]_g = -e;
- fmt_switch_forward keeps every member's message: decrypting member i of the
  batch under s'_i gives the decryption of input i under s, up to small noise.
"""
import numpy as np
import pytest

BITS = [50, 40, 40, 40] + [50, 50]   # 4 data primes at N = 2^10, 2 special primes
NP = 2


@pytest.fixture(scope="module")
def ref(oracle_lib):
    return oracle_lib.Ref(10, BITS, NP)


def centred(v, q):
    v = int(v)
    return v - q if v > q // 2 else v


def test_extend_exact_matches_crt(ref):
    N, Ql = ref.N, 4
    qs = [int(q) for q in ref.primes[:Ql]]
    Q = 1
    for q in qs:
        Q *= q
    rng = np.random.default_rng(1)
    xs = [int(rng.integers(0, 2 ** 62)) * int(rng.integers(0, 2 ** 62)) * int(rng.integers(0, 2 ** 50)) % Q
          for _ in range(N - 8)]
    xs += [0, 1, Q - 1, Q - 2, Q // 2, Q // 2 + 1, qs[0], Q - qs[0]]
    coef = np.array([[x % q for x in xs] for q in qs], dtype=np.uint64).ravel()
    outP = ref.extend_exact(coef, Ql).reshape(NP, N)
    top = [x * 10 ** 9 > Q * (10 ** 9 - 1) for x in xs]   # guard region
    assert sum(top) == 3 and top[-6] and top[-5] and top[-1] and not top[-3]  # Q-1, Q-2, Q-q0; not Q/2+1
    for k in range(NP):
        p = int(ref.primes[ref.nQ + k])
        expect = np.array([(x - Q if t else x) % p for x, t in zip(xs, top)], dtype=np.uint64)
        assert np.array_equal(outP[k], expect)


def _aux(ref, n, hw=16):
    return [ref.sk_ntt(ref.sample_ternary(100 + i, hw)) for i in range(n)]


def test_format_key_identity(ref):
    N, Ql, n = ref.N, 4, 3
    QlP = Ql + NP
    sk = ref.sk_ntt(ref.sample_ternary(7, 16))
    aux = _aux(ref, n)
    a_fmt, B = ref.gen_fmt_key(sk, aux, Ql, seed=9)
    idx = list(range(Ql)) + [ref.nQ + k for k in range(NP)]
    P = 1
    for k in range(NP):
        P *= int(ref.primes[ref.nQ + k])
    for j in range(n):
        for i in range(n):
            b = B[(j * n + i) * QlP * N:(j * n + i + 1) * QlP * N].reshape(QlP, N)
            a = a_fmt[j * QlP * N:(j + 1) * QlP * N].reshape(QlP, N)
            for k, pi in enumerate(idx[:2] + idx[-1:]):
                kk = idx.index(pi)
                q = int(ref.primes[pi])
                s_aux = aux[i].reshape(-1, N)[pi]
                v = (b[kk].astype(object) + a[kk].astype(object) * s_aux.astype(object)) % q
                if i == j and kk < Ql:
                    v = (v - (P % q) * sk.reshape(-1, N)[pi].astype(object)) % q
                e = ref.intt(np.array(v, dtype=np.uint64), [pi])
                assert max(abs(centred(x, q)) for x in e) <= 19


def test_shared_galois_key_identity(ref):
    N, n = ref.N, 2
    aux = _aux(ref, n)
    elt = ref.elt(3)
    A, B = ref.gen_shared_galois_key(aux, elt, seed=11)
    QP, dnum = ref.nQP, ref.nQ // NP
    P = 1
    for k in range(NP):
        P *= int(ref.primes[ref.nQ + k])
    for g in range(dnum):
        Ag = A[g * QP * N:(g + 1) * QP * N].reshape(QP, N)
        for i in range(n):
            b = B[(g * n + i) * QP * N:(g * n + i + 1) * QP * N].reshape(QP, N)
            sig = ref.apply_galois(aux[i], elt, QP).reshape(QP, N)
            for pi in (0, g * NP, ref.nQ):
                q = int(ref.primes[pi])
                v = (b[pi].astype(object) + Ag[pi].astype(object) * aux[i].reshape(QP, N)[pi].astype(object)) % q
                if g * NP <= pi < (g + 1) * NP:
                    v = (v - (P % q) * sig[pi].astype(object)) % q
                e = ref.intt(np.array(v, dtype=np.uint64), [pi])
                assert max(abs(centred(x, q)) for x in e) <= 19


def test_fmt_switch_forward_keeps_messages(ref):
    # FmtSwitch needs P >= Ql (noise ext * e / P): 2 primes (90 bits) under P (100 bits)
    N, Ql, n = ref.N, 2, 3
    sk = ref.sk_ntt(ref.sample_ternary(7, 16))
    aux = _aux(ref, n)
    a_fmt, B = ref.gen_fmt_key(sk, aux, Ql, seed=9)
    rng = np.random.default_rng(3)
    cts, pts = [], []
    for j in range(n):
        pt = ref.encode(rng.uniform(-1, 1, N // 2), Ql, 2.0 ** 30)
        pts.append(pt)
        cts.append(ref.encrypt_symmetric(sk, pt, Ql, 20 + j))
    out = ref.fmt_switch_forward(cts, Ql, a_fmt, B).reshape(n + 1, Ql * N)
    q0 = int(ref.primes[0])
    for i in range(n):
        member = np.concatenate([out[i], out[n]])
        d_new = ref.decrypt(aux[i], member, 2, Ql)
        d_old = ref.decrypt(sk, cts[i], 2, Ql)
        diff = (d_new[:N].astype(object) - d_old[:N].astype(object)) % q0
        e = ref.intt(np.array(diff, dtype=np.uint64), [0])
        assert max(abs(centred(x, q0)) for x in e) < 2 ** 12
        # and the message itself survives
        dm = ref.decode(d_new, Ql, 2.0 ** 30)[0]
        assert np.max(np.abs(dm - ref.decode(pts[i], Ql, 2.0 ** 30)[0])) < 1e-4
