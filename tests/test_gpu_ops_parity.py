"""GPU parity of the evaluator operations not covered elsewhere, bit-exact against
the oracle through the C ABI:

- sub, negate, add_plain, sub_plain, double (evaluate.cu:104-348, 1114-1235;
  ckks_evaluator_parallel.cuh:527-531), at N = 2^13 (8 data primes) and N = 2^16;
- mod_switch_to_next / mod_switch_to for ciphertexts and plaintexts
  (evaluate.cu:1445-1561);
- square (polymath.cu:505-550) against the tensor product of a ciphertext with itself;
- complex_conjugate (evaluate.cu:1717-1725) and the NAF split of rotate_internal
  for steps without their own key (evaluate.cu:1650-1706, numth.h:17-34);
- public-key generation and encrypt_asymmetric (secretkey.cu:15-217, 524-535);
- the facade's constant and reduced-error operations
  (ckks_evaluator_parallel.cuh:444-561, ckks_evaluator_parallel.cu:466-590);
- batch_input (Batch_encode_encrypt.cuh:12-43);
- the reference's own known answer pi x^3 + 0.4 x + 1 (test_phantom_ckks.cuh),
  run through public-key encryption, square, relinearise, rescale, multiply_plain,
  mod switch and add_plain.
"""
import numpy as np
import pytest

import shmoai_amd as shm

pytestmark = pytest.mark.gpu

C1_BITS = [50] * 8 + [60]


def _rand_limbs(ref, rng, npolys, Ql):
    out = np.empty((npolys, Ql, ref.N), np.uint64)
    for k in range(Ql):
        out[:, k, :] = rng.integers(0, int(ref.primes[k]), size=(npolys, ref.N), dtype=np.uint64)
    return out.ravel()


@pytest.fixture(scope="module", params=[(13, C1_BITS, 8), (16, [50] * 3 + [60], 3)], ids=["N2^13", "N2^16"])
def env(request, oracle_lib):
    logN, bits, Ql = request.param
    return shm.Context(logN, bits, 1), oracle_lib.Ref(logN, bits, 1), Ql


def test_elementwise_ops_bitexact(env):
    ctx, ref, Ql = env
    rng = np.random.default_rng(5)
    x, y = _rand_limbs(ref, rng, 2, Ql), _rand_limbs(ref, rng, 2, Ql)
    p = _rand_limbs(ref, rng, 1, Ql)
    ev = shm.Evaluator(ctx)
    chain = ctx.chain(Ql)

    def ct_of(v):
        c = shm.Ciphertext(ctx, 2, chain).load(v)
        c.scale = 2.0 ** 30
        return c

    pt = shm.Plaintext(ctx, chain).load(p)
    pt.scale = 2.0 ** 30
    a = ct_of(x)
    ev.sub_inplace(a, ct_of(y))
    assert np.array_equal(a.numpy(), ref.sub(x, y, Ql, 2))
    a = ct_of(x)
    ev.negate_inplace(a)
    assert np.array_equal(a.numpy(), ref.negate(x, Ql, 2))
    S = Ql * ref.N
    a = ct_of(x)
    ev.add_plain_inplace(a, pt)
    assert np.array_equal(a.numpy(), np.concatenate([ref.add(x[:S], p, Ql, 1), x[S:]]))
    a = ct_of(x)
    ev.sub_plain_inplace(a, pt)
    assert np.array_equal(a.numpy(), np.concatenate([ref.sub(x[:S], p, Ql, 1), x[S:]]))
    a = ct_of(x)
    ev.double_inplace(a)
    assert np.array_equal(a.numpy(), ref.add(x, x, Ql, 2))
    # the reference's sub checks the scales (evaluate.cu:276-277); add does not
    b = ct_of(y)
    b.scale = 2.0 ** 31
    with pytest.raises(shm.capi.ShmInvalidArgument):
        ev.sub_inplace(ct_of(x), b)


def test_mod_switch_bitexact(env):
    ctx, ref, Ql = env
    rng = np.random.default_rng(6)
    x = _rand_limbs(ref, rng, 2, Ql)
    p = _rand_limbs(ref, rng, 1, Ql)
    ev = shm.Evaluator(ctx)
    chain = ctx.chain(Ql)
    a = shm.Ciphertext(ctx, 2, chain).load(x)
    ev.mod_switch_to_next_inplace(a)
    assert a.chain_index == chain + 1
    exp = ref.mod_switch_drop(x, Ql)
    assert np.array_equal(a.numpy(), exp)
    a = shm.Ciphertext(ctx, 2, chain).load(x)
    ev.mod_switch_to_inplace(a, chain + 2)
    assert np.array_equal(a.numpy(), ref.mod_switch_drop(exp, Ql - 1))
    pt = shm.Plaintext(ctx, chain).load(p)
    ev.plain_mod_switch_to_inplace(pt, chain + 2)
    assert np.array_equal(pt.numpy(), p[:(Ql - 2) * ref.N])
    with pytest.raises(shm.capi.ShmInvalidArgument):
        ev.mod_switch_to_inplace(a, chain)  # to a higher level


def test_square_bitexact(env):
    ctx, ref, Ql = env
    rng = np.random.default_rng(7)
    x = _rand_limbs(ref, rng, 2, Ql)
    a = shm.Ciphertext(ctx, 2, ctx.chain(Ql)).load(x)
    a.scale = 2.0 ** 20
    shm.Evaluator(ctx).square_inplace(a)
    assert a.size == 3 and a.scale == 2.0 ** 40
    assert np.array_equal(a.numpy(), ref.tensor(x, x, Ql))


def test_conjugate_and_naf_rotation_bitexact(oracle_lib):
    ctx = shm.Context(13, C1_BITS, 1)
    ref = oracle_lib.Ref(13, C1_BITS, 1)
    Ql, N = 8, ref.N
    rng = np.random.default_rng(8)
    x = _rand_limbs(ref, rng, 2, Ql)
    sk = shm.SecretKey(ctx, seed=1, hamming_weight=64)
    s_ref = ref.sk_ntt(ref.sample_ternary(1, 64))
    steps = [0, -1, 4, 8]
    gk = sk.create_galois_keys_from_steps(steps, seed=9)
    keys = {st: ref.gen_galois_key(s_ref, ref.elt(st), ref.splitmix64(9, ref.elt(st))) for st in steps}
    ev = shm.Evaluator(ctx)
    # complex_conjugate = apply_galois with 2N - 1 (the key of "step 0")
    assert ref.elt(0) == 2 * N - 1
    a = shm.Ciphertext(ctx, 2, 1).load(x)
    ev.complex_conjugate_inplace(a, gk)
    assert np.array_equal(a.numpy(), ref.rotate(x, Ql, 2 * N - 1, keys[0]))
    # steps 3 = 4 - 1 and 7 = 8 - 1 have no key: NAF digits in increasing order
    for step, naf in ((3, [-1, 4]), (7, [-1, 8])):
        a = shm.Ciphertext(ctx, 2, 1).load(x)
        ev.rotate_vector_inplace(a, step, gk)
        exp = x
        for t in naf:
            exp = ref.rotate(exp, Ql, ref.elt(t), keys[t])
        assert np.array_equal(a.numpy(), exp), f"step {step}"
    # a power of two without its key cannot be split further
    with pytest.raises(shm.capi.ShmInvalidArgument):
        ev.rotate_vector_inplace(shm.Ciphertext(ctx, 2, 1).load(x), 2, gk)


@pytest.mark.parametrize("logN,bits,nP", [(13, C1_BITS, 1), (13, [50] * 6 + [60] * 3, 3)], ids=["P1", "P3"])
def test_public_key_encrypt_asymmetric_bitexact(oracle_lib, logN, bits, nP):
    ctx = shm.Context(logN, bits, nP)
    ref = oracle_lib.Ref(logN, bits, nP)
    nQ, N = ref.nQ, ref.N
    sk = shm.SecretKey(ctx, seed=3, hamming_weight=64)
    s_ref = ref.sk_ntt(ref.sample_ternary(3, 64))
    pk = sk.gen_publickey(seed=17)
    pk_ref = ref.gen_public_key(s_ref, 17)
    assert np.array_equal(pk.numpy(), pk_ref)
    # the key encrypts zero under s at the key level: pk0 + pk1 s is the (small) error
    v = np.random.default_rng(9).uniform(-1, 1, N // 2)
    enc = shm.Encoder(ctx)
    pt = enc.encode(v, 2.0 ** 40, 1)
    ct = shm.Encryptor(ctx, pk).encrypt(pt, seed=23)
    exp = ref.encrypt_asymmetric(pk_ref, ref.encode(v, nQ, 2.0 ** 40), 23)
    assert np.array_equal(ct.numpy(), exp)
    assert ct.scale == 2.0 ** 40 and ct.chain_index == 1
    got = enc.decode(shm.Decryptor(ctx, sk).decrypt(ct))
    assert float(np.max(np.abs(got - v))) < 1e-6
    low = enc.encode(v, 2.0 ** 40, 2)
    with pytest.raises(shm.capi.ShmInvalidArgument):
        shm.Encryptor(ctx, pk).encrypt(low, seed=1)


@pytest.fixture(scope="module")
def c1(oracle_lib):
    ctx = shm.Context(13, C1_BITS, 1)
    ref = oracle_lib.Ref(13, C1_BITS, 1)
    sk = shm.SecretKey(ctx, seed=1, hamming_weight=64)
    s_ref = ref.sk_ntt(ref.sample_ternary(1, 64))
    return ctx, ref, sk, s_ref


def _encrypt_pair(ctx, ref, sk, s_ref, v, Ql, scale, seed):
    enc = shm.Encoder(ctx)
    pt = enc.encode(v, scale, ctx.chain(Ql))
    ct = sk.encrypt_symmetric(pt, seed=seed)
    return ct, ref.encrypt_symmetric(s_ref, ref.encode(v, Ql, scale), Ql, seed)


def test_facade_constant_ops_bitexact(c1):
    ctx, ref, sk, s_ref = c1
    n, N, Ql, scale = ctx.N // 2, ctx.N, 6, 2.0 ** 40
    rng = np.random.default_rng(10)
    v = rng.uniform(-1, 1, n)
    ev = shm.Evaluator(ctx)
    S = Ql * N
    # multiply_const: constant encoded over all slots at ct.scale (real and complex)
    for c in (0.731, -2.5, complex(0.25, -0.75)):
        ct, cr = _encrypt_pair(ctx, ref, sk, s_ref, v, Ql, scale, 31)
        ev.multiply_const_inplace(ct, c)
        if isinstance(c, complex):
            pt = ref.encode(np.full(n, c.real), Ql, scale, im=np.full(n, c.imag))
        else:
            pt = ref.encode(np.full(n, c), Ql, scale)
        assert np.array_equal(ct.numpy(), ref.multiply_plain(cr, pt, Ql)), c
        assert ct.scale == scale * scale
    # add_const
    ct, cr = _encrypt_pair(ctx, ref, sk, s_ref, v, Ql, scale, 32)
    ev.add_const_inplace(ct, 0.125)
    exp = np.concatenate([ref.add(cr[:S], ref.encode(np.full(n, 0.125), Ql, scale), Ql, 1), cr[S:]])
    assert np.array_equal(ct.numpy(), exp)
    # multiply_vector_inplace_reduced_error: a short vector padded with zeros
    w = rng.uniform(-2, 2, 1000)
    ct, cr = _encrypt_pair(ctx, ref, sk, s_ref, v, Ql, scale, 33)
    ev.multiply_vector_inplace_reduced_error(ct, w)
    assert np.array_equal(ct.numpy(), ref.multiply_plain(cr, ref.encode(w, Ql, scale), Ql))
    dec = shm.Encoder(ctx).decode(shm.Decryptor(ctx, sk).decrypt(ct))
    want = np.concatenate([w, np.zeros(n - w.size)]) * v
    assert float(np.max(np.abs(dec - want))) < 1e-5


@pytest.mark.parametrize("op", ["add", "sub", "multiply"])
@pytest.mark.parametrize("levels", [(6, 6), (4, 7), (7, 4)], ids=["same", "ct1_lower", "ct1_higher"])
def test_reduced_error_ops_bitexact(c1, op, levels):
    """ckks_evaluator_parallel.cu:466-590: the operand with more limbs is multiplied
    by s_low q_last / s_high^2, labelled s_low q_last, rescaled (floor) and
    mod-switched before the plain operation; multiply relinearises."""
    ctx, ref, sk, s_ref = c1
    n, N = ctx.N // 2, ctx.N
    Q1, Q2 = levels
    # operands at one level only take over the other's scale, so they share it here
    s1, s2 = 2.0 ** 40, (2.0 ** 40 if Q1 == Q2 else 2.0 ** 38)
    rng = np.random.default_rng(11)
    v1, v2 = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    ct1, r1 = _encrypt_pair(ctx, ref, sk, s_ref, v1, Q1, s1, 41)
    ct2, r2 = _encrypt_pair(ctx, ref, sk, s_ref, v2, Q2, s2, 42)
    rk = sk.gen_relinkey(seed=5)
    rk_ref = ref.gen_relin_key(s_ref, 5)
    ev = shm.Evaluator(ctx)

    def adjust(rc, Qh, sh, Ql_low, s_low):
        qlast = float(ref.primes[Qh - 1])
        k = s_low * qlast / (sh * sh)
        t = ref.rescale(ref.multiply_plain(rc, ref.encode(np.full(n, k), Qh, sh), Qh), Qh)
        return t.reshape(2, Qh - 1, N)[:, :Ql_low].ravel()

    def plain(a, b, Q):
        if op == "add":
            return ref.add(a, b, Q, 2)
        if op == "sub":
            return ref.sub(a, b, Q, 2)
        return ref.relinearize(ref.tensor(a, b, Q), Q, rk_ref)

    if Q1 == Q2:
        exp, Q, want_scale = plain(r1, r2, Q1), Q1, s2 * (s2 if op == "multiply" else 1.0)
    elif Q1 < Q2:
        exp, Q = plain(r1, adjust(r2, Q2, s2, Q1, s1), Q1), Q1
        want_scale = s1 * (s1 if op == "multiply" else 1.0)
    else:
        exp, Q = plain(adjust(r1, Q1, s1, Q2, s2), r2, Q2), Q2
        want_scale = s2 * (s2 if op == "multiply" else 1.0)
    getattr(ev, f"{op}_inplace_reduced_error")(*((ct1, ct2, rk) if op == "multiply" else (ct1, ct2)))
    assert ct1.chain_index == ctx.chain(Q) and ct1.size == 2
    assert np.array_equal(ct1.numpy(), exp)
    assert ct1.scale == pytest.approx(want_scale, rel=1e-12)
    got = shm.Encoder(ctx).decode(shm.Decryptor(ctx, sk).decrypt(ct1))
    want = {"add": v1 + v2, "sub": v1 - v2, "multiply": v1 * v2}[op]
    assert float(np.max(np.abs(got - want))) < 1e-4


def test_batch_input_bitexact(c1):
    ctx, ref, sk, s_ref = c1
    n, nQ = ctx.N // 2, ref.nQ
    num_x, num_row, num_col, scale = 4, 16, 5, 2.0 ** 40
    rng = np.random.default_rng(12)
    X = rng.normal(0, 1, (num_x, num_row, num_col))
    pk = sk.gen_publickey(seed=19)
    pk_ref = ref.gen_public_key(s_ref, 19)
    cts = shm.batch_input(X, scale, ctx, pk, seed=77)
    assert len(cts) == num_col
    enc, dec = shm.Encoder(ctx), shm.Decryptor(ctx, sk)
    for i in range(num_col):
        vec = np.zeros(n)
        for j in range(num_x):
            for k in range(num_row):
                vec[num_x * k + j] = X[j][k][i]
        exp = ref.encrypt_asymmetric(pk_ref, ref.encode(vec, nQ, scale), ref.splitmix64(77, i))
        assert np.array_equal(cts[i].numpy(), exp), f"column {i}"
        assert float(np.max(np.abs(enc.decode(dec.decrypt(cts[i])) - vec))) < 1e-6
    with pytest.raises(shm.capi.ShmInvalidArgument):
        shm.batch_input(np.zeros((n, 2, 1)), scale, ctx, pk, seed=1)  # num_X * num_row > N/2


def test_known_answer_cubic_gpu(oracle_lib):
    """test_phantom_ckks.cuh: N = 2^15, primes {60, 40, 40, 60}, scale 2^40, the input
    accumulated in steps of 1/(slots - 1), pi x^3 + 0.4 x + 1 with the reference's
    sequence of calls; the ciphertext equals the oracle's composition limb for limb
    and decrypts to the polynomial."""
    bits = [60, 40, 40, 60]
    ctx = shm.Context(15, bits, 1)
    ref = oracle_lib.Ref(15, bits, 1)
    n, scale = ctx.N // 2, 2.0 ** 40
    x = np.empty(n)
    cur, step = 0.0, 1.0 / (float(n) - 1.0)
    for i in range(n):
        x[i] = cur
        cur += step
    sk = shm.SecretKey(ctx, seed=51, hamming_weight=64)
    s_ref = ref.sk_ntt(ref.sample_ternary(51, 64))
    pk, rk = sk.gen_publickey(seed=52), sk.gen_relinkey(seed=53)
    pk_ref, rk_ref = ref.gen_public_key(s_ref, 52), ref.gen_relin_key(s_ref, 53)
    enc, ev = shm.Encoder(ctx), shm.Evaluator(ctx)
    c3, c1_, c0 = (enc.encode(v, scale, 1) for v in (3.14159265, 0.4, 1.0))
    x1 = shm.Encryptor(ctx, pk).encrypt(enc.encode(x, scale, 1), seed=54)
    x3 = x1.copy()
    ev.square_inplace(x3)
    ev.relinearize_inplace(x3, rk)
    ev.rescale_to_next_inplace(x3)
    x1c3 = x1.copy()
    ev.multiply_plain_inplace(x1c3, c3)
    ev.rescale_to_next_inplace(x1c3)
    ev.multiply_inplace(x3, x1c3)
    ev.relinearize_inplace(x3, rk)
    ev.rescale_to_next_inplace(x3)
    ev.multiply_plain_inplace(x1, c1_)
    ev.rescale_to_next_inplace(x1)
    x3.scale = scale
    x1.scale = scale
    ev.mod_switch_to_inplace(x1, x3.chain_index)
    ev.plain_mod_switch_to_inplace(c0, x3.chain_index)
    res = x3.copy()
    ev.add_inplace(res, x1)
    ev.add_plain_inplace(res, c0)
    # oracle: the reference's encode(double) is the full-slot encode of a constant vector
    Q, N = 3, ref.N
    e = {v: ref.encode(np.full(n, v), Q, scale) for v in (3.14159265, 0.4, 1.0)}
    r1 = ref.encrypt_asymmetric(pk_ref, ref.encode(x, Q, scale), 54)
    r3 = ref.rescale(ref.relinearize(ref.tensor(r1, r1, Q), Q, rk_ref), Q)
    rc3 = ref.rescale(ref.multiply_plain(r1, e[3.14159265], Q), Q)
    r3 = ref.rescale(ref.relinearize(ref.tensor(r3, rc3, Q - 1), Q - 1, rk_ref), Q - 1)
    rl = ref.rescale(ref.multiply_plain(r1, e[0.4], Q), Q)
    rl = ref.mod_switch_drop(rl, Q - 1)
    rr = ref.add(r3, rl, Q - 2, 2)
    rr[:(Q - 2) * N] = ref.add(rr[:(Q - 2) * N], e[1.0][:(Q - 2) * N], Q - 2, 1)
    assert np.array_equal(res.numpy(), rr)
    got = enc.decode(shm.Decryptor(ctx, sk).decrypt(res))
    want = (3.14159265 * x * x + 0.4) * x + 1
    assert float(np.max(np.abs(got - want))) < 1e-3
