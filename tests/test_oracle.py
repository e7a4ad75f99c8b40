"""Pin the oracle (oracle/refcpu) before trusting it as the GPU checker.

Each check compares the oracle with an independent definition (tests/_pyref.py,
Python big integers) or a mathematical identity that a plausible mistake would
break: brute-force prime enumeration, brute-force minimal roots, direct
polynomial evaluation for the NTT, schoolbook negacyclic products, coefficient
domain automorphisms, CRT floor division for rescale, the exact fast base
conversion identity for ModUp, ModDown of P*y, decryption noise bounds for key
switching, and the reference's decoded known answers.
"""
import json
import os

import numpy as np
import pytest

from tests import _pyref as P

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def small(oracle_lib):
    return oracle_lib.Ref(6, [20, 21, 22, 23], 1)  # N = 64


@pytest.mark.parametrize("logN,bits,nP", [
    (13, [50] * 8 + [60], 1),
    (16, [50] * 24 + [60], 1),
    (16, [51] + [46] * 20 + [51] * 14 + [51] * 5, 5),
    (16, [60] + [40] * 14 + [60], 1),
])
def test_prime_chain_matches_bruteforce(oracle_lib, logN, bits, nP):
    r = oracle_lib.Ref(logN, bits, nP)
    exp = P.coeff_modulus(1 << logN, bits)
    assert [int(x) for x in r.primes] == exp
    for q, b in zip(exp, bits):
        assert q % (2 << logN) == 1 and (1 << (b - 1)) < q < (1 << b) and P.is_prime(q)


def test_minimal_root_bruteforce(oracle_lib):
    r = oracle_lib.Ref(3, [17, 18, 19], 1)  # N = 8, q ~ 2^17..2^19
    for q, psi in zip(r.primes, r.psi()):
        q, psi = int(q), int(psi)
        x = 2
        while pow(x, 8, q) != q - 1:
            x += 1
        assert psi == x


def test_minimal_root_properties_c1(oracle_lib):
    r = oracle_lib.Ref(13, [50] * 8 + [60], 1)
    N = r.N
    for q, psi in zip(r.primes[:3], r.psi()[:3]):
        q, psi = int(q), int(psi)
        assert pow(psi, N, q) == q - 1
        g2, cur, best = psi * psi % q, psi, psi
        for _ in range(N):
            best = min(best, cur)
            cur = cur * g2 % q
        assert best == psi


def test_ntt_matches_direct_evaluation(small):
    rng = np.random.default_rng(1)
    for k in range(4):
        q, psi = int(small.primes[k]), int(small.psi()[k])
        a = rng.integers(0, q, small.N, dtype=np.uint64)
        got = small.ntt(a, np.array([k], np.uint32))
        assert [int(v) for v in got] == P.ntt_direct([int(v) for v in a], psi, q)


def test_intt_inverse_and_negacyclic_convolution(small):
    rng = np.random.default_rng(2)
    k = 1
    q = int(small.primes[k])
    idx = np.array([k], np.uint32)
    a = rng.integers(0, q, small.N, dtype=np.uint64)
    b = rng.integers(0, q, small.N, dtype=np.uint64)
    A, B = small.ntt(a, idx), small.ntt(b, idx)
    assert np.array_equal(small.intt(A, idx), a)
    prod = np.array([int(x) * int(y) % q for x, y in zip(A, B)], np.uint64)
    assert [int(v) for v in small.intt(prod, idx)] == P.negacyclic_mul([int(v) for v in a], [int(v) for v in b], q)


@pytest.mark.parametrize("step", [1, -1, 3, 7, 0])
def test_galois_permutation_is_automorphism(small, step):
    rng = np.random.default_rng(3)
    elt = small.elt(step)
    assert elt == (2 * small.N - 1 if step == 0 else pow(5, step % (small.N // 2), 2 * small.N))
    k = 2
    q, psi = int(small.primes[k]), int(small.psi()[k])
    a = [int(v) for v in rng.integers(0, q, small.N, dtype=np.uint64)]
    A = np.array(P.ntt_direct(a, psi, q), np.uint64)
    got = small.apply_galois(A, elt, 1)
    assert [int(v) for v in got] == P.ntt_direct(P.automorphism(a, elt, q), psi, q)


def test_rescale_is_floor_division(small):
    rng = np.random.default_rng(4)
    Ql = 3
    primes = [int(q) for q in small.primes[:Ql]]
    coef = np.stack([rng.integers(0, q, small.N, dtype=np.uint64) for q in primes])
    ntt = small.ntt(coef.ravel(), small.pidx(Ql))
    out = small.rescale(ntt, Ql, npolys=1)
    oc = small.intt(out, small.pidx(Ql - 1)).reshape(Ql - 1, small.N)
    for j in range(small.N):
        C = P.crt([int(coef[i, j]) for i in range(Ql)], primes)
        f = C // primes[-1]
        assert [int(oc[i, j]) for i in range(Ql - 1)] == [f % p for p in primes[:-1]]


@pytest.mark.parametrize("nP", [1, 2])
def test_modup_is_fast_base_conversion(oracle_lib, nP):
    bits = [30, 31, 32, 33, 34] + [40] * nP
    r = oracle_lib.Ref(6, bits, nP)
    rng = np.random.default_rng(5)
    Ql = 5
    primes = [int(q) for q in r.primes]
    coef = np.stack([rng.integers(0, primes[i], r.N, dtype=np.uint64) for i in range(Ql)])
    c2 = r.ntt(coef.ravel(), r.pidx(Ql))
    d = r.modup(c2, Ql)
    QlP = Ql + nP
    pidx = list(range(Ql)) + list(range(r.nQ, r.nQ + nP))
    beta = r.beta(Ql)
    d = d.reshape(beta, QlP, r.N)
    for g in range(beta):
        start, size = g * nP, min(nP, Ql - g * nP)
        digit = primes[start:start + size]
        D = int(np.prod([x for x in digit], dtype=object))
        # own limbs are copied in NTT form
        assert np.array_equal(d[g, start:start + size], c2.reshape(Ql, r.N)[start:start + size])
        dc = r.intt(d[g].ravel(), np.array(pidx, np.uint32)).reshape(QlP, r.N)
        for j in range(0, r.N, 7):
            xs = [int(coef[start + i, j]) for i in range(size)]
            X = P.crt(xs, digit)
            V = sum((x * pow(D // qd, -1, qd) % qd) * (D // qd) for x, qd in zip(xs, digit))
            u, rem = divmod(V - X, D)
            assert rem == 0 and 0 <= u < size
            for k in range(QlP):
                if start <= k < start + size:
                    continue
                assert int(dc[k, j]) == V % primes[pidx[k]]


@pytest.mark.parametrize("nP", [1, 3])
def test_moddown_of_p_times_y_is_y(oracle_lib, nP):
    bits = [30, 31, 32, 33] + [40] * nP
    r = oracle_lib.Ref(6, bits, nP)
    rng = np.random.default_rng(6)
    Ql = 4
    primes = [int(q) for q in r.primes]
    Pbig = int(np.prod([primes[r.nQ + j] for j in range(nP)], dtype=object))
    y = np.stack([rng.integers(0, primes[i], r.N, dtype=np.uint64) for i in range(Ql)])
    pidx = list(range(Ql)) + list(range(r.nQ, r.nQ + nP))
    cx = np.zeros((Ql + nP, r.N), np.uint64)
    for k in range(Ql):
        cx[k] = [int(v) * Pbig % primes[k] for v in y[k]]
    cxn = r.ntt(cx.ravel(), np.array(pidx, np.uint32))
    out = r.moddown(cxn, Ql)
    assert np.array_equal(r.intt(out, r.pidx(Ql)).reshape(Ql, r.N), y)


def test_moddown_congruence(oracle_lib):
    bits = [30, 31, 32] + [40, 41]
    r = oracle_lib.Ref(6, bits, 2)
    rng = np.random.default_rng(7)
    Ql, nP = 3, 2
    primes = [int(q) for q in r.primes]
    pidx = list(range(Ql)) + [r.nQ, r.nQ + 1]
    cxc = np.stack([rng.integers(0, primes[p], r.N, dtype=np.uint64) for p in pidx])
    cx = r.ntt(cxc.ravel(), np.array(pidx, np.uint32))
    out = r.intt(r.moddown(cx, Ql), r.pidx(Ql)).reshape(Ql, r.N)
    Pm = [primes[r.nQ], primes[r.nQ + 1]]
    Pbig = Pm[0] * Pm[1]
    for j in range(0, r.N, 5):
        xp = [int(cxc[Ql + t, j]) for t in range(nP)]
        delta = sum((x * pow(Pbig // p, -1, p) % p) * (Pbig // p) for x, p in zip(xp, Pm))
        for i in range(Ql):
            q = primes[i]
            assert int(out[i, j]) * Pbig % q == (int(cxc[i, j]) - delta) % q


def _centered(v, Q):
    return v - Q if v > Q // 2 else v


@pytest.mark.parametrize("bits,nP", [([50] * 4 + [60], 1), ([50] * 4 + [60, 60], 2)])
def test_keyswitch_decrypts_within_noise(oracle_lib, bits, nP):
    r = oracle_lib.Ref(10, bits, nP)
    Ql = r.nQ
    s = r.sk_ntt(r.sample_ternary(11, 64))
    s2 = r.sk_ntt(r.sample_ternary(12, 64))
    key = r.gen_kswitch_key(s, s2, 13)
    rng = np.random.default_rng(8)
    c2 = np.concatenate([rng.integers(0, int(r.primes[i]), r.N, dtype=np.uint64) for i in range(Ql)])
    k0, k1 = r.keyswitch(c2, key, Ql)
    primes = [int(q) for q in r.primes[:Ql]]
    S = Ql * r.N
    # ks0 + ks1*s - c2*s2 (NTT domain), to coefficients
    diff = np.zeros(S, np.uint64)
    for i, q in enumerate(primes):
        sl = slice(i * r.N, (i + 1) * r.N)
        a = (k0[sl].astype(object) + k1[sl].astype(object) * s[sl].astype(object)
             - c2[sl].astype(object) * s2[sl].astype(object)) % q
        diff[sl] = np.array(a, np.uint64)
    dc = r.intt(diff, r.pidx(Ql)).reshape(Ql, r.N)
    Q = int(np.prod(primes, dtype=object))
    noise = max(abs(_centered(P.crt([int(dc[i, j]) for i in range(Ql)], primes), Q)) for j in range(r.N))
    assert 0 < noise < 2 ** 20


def _decrypt_noise(r, s_ntt, ct, pt, Ql, limbs_idx=None):
    """max |c0 + c1 s - m| over the coefficients (centred CRT value)."""
    N = r.N
    idx = r.pidx(Ql) if limbs_idx is None else limbs_idx
    primes = [int(r.primes[k]) for k in idx]
    L = len(primes)
    diff = np.zeros(L * N, np.uint64)
    for i, q in enumerate(primes):
        sl = slice(i * N, (i + 1) * N)
        a = (ct[sl].astype(object) + ct[L * N:][sl].astype(object) * s_ntt[idx[i] * N:(idx[i] + 1) * N].astype(object)
             - pt[sl].astype(object)) % q
        diff[sl] = np.array(a, np.uint64)
    dc = r.intt(diff, np.array(idx, np.uint32)).reshape(L, N)
    Q = int(np.prod(primes, dtype=object))
    return max(abs(_centered(P.crt([int(dc[i, j]) for i in range(L)], primes), Q)) for j in range(N))


@pytest.mark.parametrize("bits,nP", [([50] * 4 + [60], 1), ([50] * 4 + [60, 60], 2)])
def test_public_key_and_asymmetric_encryption_decrypt(oracle_lib, bits, nP):
    """gen_publickey is an encryption of zero at the key level: pk0 + pk1 s = -e
    with the Gaussian error (|e| <= 19); encrypt_asymmetric decrypts to the
    plaintext plus the noise (u e_pk + e + e s) / P + ModDown rounding, far below
    2^20 (a dropped ModDown, a wrong sign or a missing u would leave noise of the
    size of the modulus)."""
    r = oracle_lib.Ref(10, bits, nP)
    s = r.sk_ntt(r.sample_ternary(61, 64))
    pk = r.gen_public_key(s, 62)
    S = r.nQP * r.N
    zero = np.zeros(S, np.uint64)
    assert 0 < _decrypt_noise(r, s, pk, zero, r.nQP, list(range(r.nQP))) <= 19
    u = r.sample_ternary_uniform(5)
    assert set(np.unique(u)) == {-1, 0, 1} and abs(np.count_nonzero(u) / r.N - 2 / 3) < 0.05
    rng = np.random.default_rng(3)
    Ql = r.nQ
    pt = np.concatenate([rng.integers(0, int(r.primes[i]), r.N, dtype=np.uint64) for i in range(Ql)])
    ct = r.encrypt_asymmetric(pk, pt, 63)
    noise = _decrypt_noise(r, s, ct, pt, Ql)
    assert 0 < noise < 2 ** 20


def test_samplers(oracle_lib):
    r = oracle_lib.Ref(14, [50, 50, 60], 1)
    t = r.sample_ternary(5, 192)
    assert np.count_nonzero(t) == 192 and set(np.unique(t)) <= {-1, 0, 1}
    e = np.concatenate([r.sample_gaussian(s) for s in range(8)])
    assert np.abs(e).max() <= 19 and abs(e.mean()) < 0.05 and abs(e.std() - 3.2) < 0.05
    u = r.sample_uniform(9, r.pidx(2))
    q0 = int(r.primes[0])
    assert int(u[:r.N].max()) < q0 and abs(u[:r.N].astype(np.float64).mean() / q0 - 0.5) < 0.01
    assert r.splitmix64(0, 0) == 0xE220A8397B1DCDAF  # SplitMix64 reference value for state 0


def test_scalar_encode_equals_full_fft_path(oracle_lib):
    r = oracle_lib.Ref(12, [60, 40, 40, 60], 1)
    for w in (0.0123456789, -0.02, 1.0 / 128.0, -3.5e-7):
        got = r.encode_scalar(w, 3, 2.0 ** 40)
        full = r.encode(np.full(r.N // 2, w), 3, 2.0 ** 40)
        assert np.array_equal(got, full)
        v = int(np.round(w * 2.0 ** 40))
        for k in range(3):
            assert np.all(got[k * r.N:(k + 1) * r.N] == v % int(r.primes[k]))


def test_encode_matches_vandermonde_solution(oracle_lib):
    """The encoder's coefficients against the canonical embedding solved directly in
    numpy (no FFT): m(w_j) = scale z_j and m(conj w_j) = scale conj(z_j) with
    w_j = exp(2 pi i 5^j / 2N), j < N/2, a linear system for the N real coefficients;
    the oracle's plaintext, brought back to coefficients, equals round(m) per prime."""
    r = oracle_lib.Ref(6, [40, 41, 50], 1)  # N = 64, two data primes
    N, n = r.N, r.N // 2
    rng = np.random.default_rng(10)
    z = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    scale = 2.0 ** 30
    roots = np.exp(2j * np.pi * np.array([pow(5, j, 2 * N) for j in range(n)]) / (2 * N))
    V = np.vstack([roots[:, None] ** np.arange(N), np.conj(roots)[:, None] ** np.arange(N)])
    m = np.linalg.solve(V, scale * np.concatenate([z, np.conj(z)])).real
    pt = r.encode(z.real, 2, scale, im=z.imag)
    coef = r.intt(pt, r.pidx(2)).reshape(2, N)
    for k in range(2):
        q = int(r.primes[k])
        want = np.array([int(v) % q for v in np.round(m)], np.uint64)
        assert np.array_equal(coef[k], want)


def test_encode_decode_roundtrip(oracle_lib):
    r = oracle_lib.Ref(12, [60, 40, 40, 60], 1)
    rng = np.random.default_rng(9)
    v = rng.uniform(-1, 1, r.N // 2) + 1j * rng.uniform(-1, 1, r.N // 2)
    pt = r.encode(v.real, 3, 2.0 ** 40, im=v.imag)
    re, im = r.decode(pt, 3, 2.0 ** 40)
    assert np.max(np.abs(re + 1j * im - v)) < 1e-9


def test_known_answer_cubic(oracle_lib):
    """test_phantom_ckks.cuh: pi*x^3 + 0.4*x + 1 on x = linspace(0, 1)."""
    r = oracle_lib.Ref(13, [60, 40, 40, 40, 60], 1)
    s = r.sk_ntt(r.sample_ternary(21, 64))
    rk = r.gen_relin_key(s, 22)
    n = r.N // 2
    x = np.linspace(0.0, 1.0, n)
    scale = 2.0 ** 40
    Ql = 4
    ct = r.encrypt_symmetric(s, r.encode(x, Ql, scale), Ql, 23)
    x2 = r.rescale(r.relinearize(r.tensor(ct, ct, Ql), Ql, rk), Ql)               # x^2 at Ql-1
    q3 = float(r.primes[Ql - 1])
    sc2 = scale * scale / q3
    pix = r.rescale(r.multiply_plain(ct, r.encode_scalar(np.pi, Ql, scale), Ql), Ql)  # pi x at Ql-1
    sc_pix = scale * scale / q3
    x3 = r.rescale(r.relinearize(r.tensor(x2, pix, Ql - 1), Ql - 1, rk), Ql - 1)  # pi x^3 at Ql-2
    sc3 = sc2 * sc_pix / float(r.primes[Ql - 2])
    lin = r.rescale(r.multiply_plain(ct, r.encode_scalar(0.4, Ql, scale), Ql), Ql)
    lin = r.mod_switch_drop(lin, Ql - 1)
    # bring 0.4x (scale sc_pix at Ql-2) to sc3 by the exact ratio of scales
    dec3 = r.decode(r.decrypt(s, x3, 2, Ql - 2), Ql - 2, sc3)[0]
    decl = r.decode(r.decrypt(s, lin, 2, Ql - 2), Ql - 2, sc_pix)[0]
    res = dec3 + decl + 1.0
    assert np.max(np.abs(res - (np.pi * x ** 3 + 0.4 * x + 1))) < 1e-3


def test_matmul_known_answer_oracle(oracle_lib):
    """test_ct_pt_matrix_mul.cuh: X[i][j][k] = j+1, W = 1/128 -> (row_W/128)(j+1)."""
    r = oracle_lib.Ref(11, [60, 40, 40, 60], 1)
    s = r.sk_ntt(r.sample_ternary(31, 64))
    num_X, num_row = 8, 128
    slot = np.repeat(np.arange(1, num_row + 1, dtype=np.float64), num_X)
    row_w, col_w, Ql = 48, 3, 3
    cts = [r.encrypt_symmetric(s, r.encode(slot, Ql, 2.0 ** 40), Ql, 40 + i) for i in range(row_w)]
    W = np.full((row_w, col_w), 1.0 / 128.0)
    out = r.ctpt_matmul(cts, W, Ql, np.full(row_w, 2.0 ** 40))
    for o in out:
        v = r.decode(r.decrypt(s, o, 2, Ql - 1), Ql - 1, 2.0 ** 40)[0]
        assert np.max(np.abs(v - slot * row_w / 128.0)) < 1e-3


def test_masked_matmul_oracle_decrypts_to_masked_sums(oracle_lib):
    """The masked form (Ct_pt_matrix_mul.cuh:846-985): output i decrypts to
    sum_j W[j][i] x_j in the slots where bias_vec == 1 and to 0 elsewhere."""
    r = oracle_lib.Ref(11, [60, 40, 40, 60], 1)
    s = r.sk_ntt(r.sample_ternary(41, 64))
    rng = np.random.default_rng(42)
    n, row_w, col_w, Ql = r.N // 2, 5, 3, 3
    xs = [rng.uniform(-1, 1, n) for _ in range(row_w)]
    cts = [r.encrypt_symmetric(s, r.encode(x, Ql, 2.0 ** 40), Ql, 50 + i) for i, x in enumerate(xs)]
    W = rng.normal(0, 0.3, (row_w, col_w))
    mask = (rng.random(n) < 0.5).astype(np.int32)
    mask[:4] = [1, 0, 2, -1]  # only entries equal to 1 select
    out = r.ctpt_matmul_masked(cts, W, mask, Ql, np.full(row_w, 2.0 ** 40))
    for i, o in enumerate(out):
        v = r.decode(r.decrypt(s, o, 2, Ql - 1), Ql - 1, 2.0 ** 40)[0]
        want = np.where(mask == 1, sum(W[j, i] * xs[j] for j in range(row_w)), 0.0)
        assert np.max(np.abs(v - want)) < 1e-4


def test_ctct_matmul_oracle_decrypts_to_rotated_products(oracle_lib):
    """The column-packing ct-ct product (Ct_ct_matrix_mul.cuh:10-181): output i
    decrypts to sum_j x_j * roll(w_j, -i * num_batch) (slot s holds
    sum_j x_j[s] w_j[s + i num_batch]), at scale Delta^2 / q_last."""
    r = oracle_lib.Ref(11, [60, 40, 40, 60], 1)
    s = r.sk_ntt(r.sample_ternary(61, 64))
    rng = np.random.default_rng(62)
    n, col_x, row_x, nb, Ql, D = r.N // 2, 3, 4, 3, 3, 2.0 ** 40
    xs = [rng.uniform(-1, 1, n) for _ in range(col_x)]
    ws = [rng.uniform(-1, 1, n) for _ in range(col_x)]
    X = [r.encrypt_symmetric(s, r.encode(x, Ql, D), Ql, 70 + j) for j, x in enumerate(xs)]
    W = [r.encrypt_symmetric(s, r.encode(w, Ql, D), Ql, 80 + j) for j, w in enumerate(ws)]
    gk = {i * nb: r.gen_galois_key(s, r.elt(i * nb), r.splitmix64(9, r.elt(i * nb))) for i in range(1, row_x)}
    rk = r.gen_relin_key(s, 5)
    out = r.ctct_matmul_colpacking(X, W, Ql, row_x, nb, gk, rk)
    scale = D * D / float(r.primes[Ql - 1])
    for i, o in enumerate(out):
        v = r.decode(r.decrypt(s, o, 2, Ql - 1), Ql - 1, scale)[0]
        want = sum(xs[j] * np.roll(ws[j], -i * nb) for j in range(col_x))
        assert np.max(np.abs(v - want)) < 1e-4, i


def _diag_expected(xs, ws_i, nb):
    """Slot values of Ct_ct_matrix_mul.cuh:182-609 for one W ciphertext, from the
    plaintext vectors (rotate(v, s) = roll(v, -s))."""
    col_x = len(xs)
    g = int(np.sqrt(col_x))
    g += g * g < col_x
    b = -(-col_x // g)
    out = 0.0
    for j in range(b):
        giant = 0.0
        for k in range(g):
            if j * g + k >= col_x:
                break
            xr = xs[j * g + k] if j == 0 else np.roll(xs[j * g + k], -(col_x - j * g) * nb)
            giant = giant + np.roll(ws_i, -k * nb) * xr
        out = out + np.roll(giant, -j * g * nb)
    return out


def test_ctct_diag_oracle_decrypts_to_bsgs_sums(oracle_lib):
    """The diagonal-packing ct-ct product (baby-step giant-step,
    Ct_ct_matrix_mul.cuh:182-609), col_X = 5 (g = 3, b = 2): each output decrypts to
    the same baby/giant rotations and sums taken on the plaintext vectors."""
    r = oracle_lib.Ref(11, [60, 40, 40, 60], 1)
    s = r.sk_ntt(r.sample_ternary(63, 64))
    rng = np.random.default_rng(64)
    n, col_x, col_w, nb, Ql, D = r.N // 2, 5, 2, 2, 3, 2.0 ** 40
    xs = [rng.uniform(-1, 1, n) for _ in range(col_x)]
    ws = [rng.uniform(-1, 1, n) for _ in range(col_w)]
    X = [r.encrypt_symmetric(s, r.encode(x, Ql, D), Ql, 90 + j) for j, x in enumerate(xs)]
    W = [r.encrypt_symmetric(s, r.encode(w, Ql, D), Ql, 95 + j) for j, w in enumerate(ws)]
    steps = {st % n for st in (2, 4, 6, 4 * 2, 3 * 2)} - {0}   # babies k nb, pre-rotation (5-3) nb, giant 3 nb
    gk = {st: r.gen_galois_key(s, r.elt(st), r.splitmix64(9, r.elt(st))) for st in steps}
    rk = r.gen_relin_key(s, 5)
    out = r.ctct_matmul_diagpacking(X, W, Ql, nb, gk, rk)
    scale = D * D / float(r.primes[Ql - 1])
    for i, o in enumerate(out):
        v = r.decode(r.decrypt(s, o, 2, Ql - 1), Ql - 1, scale)[0]
        assert np.max(np.abs(v - _diag_expected(xs, ws[i], nb))) < 1e-4, i


def test_golden_fixtures(oracle_lib):
    """Committed vectors written by tests/golden/make_fixtures.py (pure Python)."""
    with open(os.path.join(GOLDEN, "fixtures.json")) as f:
        fx = json.load(f)
    for case in fx["primes"]:
        r = oracle_lib.Ref(case["logN"], case["bits"], case["nP"])
        assert [int(q) for q in r.primes] == case["primes"]
    for case in fx["ntt"]:
        r = oracle_lib.Ref(case["logN"], case["bits"], 1)
        k = case["prime_index"]
        got = r.ntt(np.array(case["input"], np.uint64), np.array([k], np.uint32))
        assert [int(v) for v in got] == case["output"]
    for case in fx["galois"]:
        r = oracle_lib.Ref(case["logN"], case["bits"], 1)
        assert [int(v) for v in r.galois_table(case["elt"])] == case["table"]
    for case in fx["rescale"]:
        r = oracle_lib.Ref(case["logN"], case["bits"], 1)
        Ql = case["Ql"]
        ntt_in = r.ntt(np.array(case["coeff_in"], np.uint64), r.pidx(Ql))
        out = r.intt(r.rescale(ntt_in, Ql, npolys=1), r.pidx(Ql - 1))
        assert [int(v) for v in out] == case["coeff_out"]


def test_oracle_refuses_missing_key_digits(oracle_lib):
    """Keys hold nQ / nP digits; a key switch needing more is refused rather than read
    past the key (nQ = 8, nP = 3: 2 digits, Ql = 7 needs 3)."""
    ref = oracle_lib.Ref(6, [51] + [46] * 7 + [51] * 3, 3)
    assert ref.dnum() == 2 and ref.beta(7) == 3 and ref.beta(6) == 2
    d = np.zeros(ref.beta(7) * 10 * ref.N, np.uint64)
    with pytest.raises(ValueError):
        ref.ks_inner(d, np.zeros(ref.key_size(), np.uint64), 7)
