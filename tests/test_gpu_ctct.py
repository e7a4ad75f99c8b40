"""ct-ct matrix products (Ct_ct_matrix_mul.cuh) on the GPU.

Column packing (10-181):
- bit-exact against the oracle composed of rotate, multiply (tensor), add,
  relinearize and rescale, for 3 x 5 operands with num_batch 3 (steps 3..12)
  and for num_batch 0 (every output is the unrotated sum);
- the outputs decrypt to sum_j x_j * roll(w_j, -i num_batch) within 1e-4;
- mismatched dimensions (the reference's check) and aliasing are refused.
Diagonal packing, baby-step giant-step (182-609):
- col_X = 5 (3 babies, 2 giants), 3 W ciphertexts, num_batch 2: bit-exact against
  the oracle, outputs labelled with X's scale, decrypting to the same baby/giant
  rotations and sums taken on the plaintext vectors.
Above 16 key digits (18 here) the rotations and relinearisations take single key
switches instead of the batched one: still bit-exact.
"""
import numpy as np
import pytest

import shmoai_amd as shm

pytestmark = pytest.mark.gpu

BITS = [60, 40, 40, 40, 60]
D = 2.0 ** 40


@pytest.fixture(scope="module")
def env(oracle_lib):
    ctx = shm.Context(12, BITS, 1)
    ref = oracle_lib.Ref(12, BITS, 1)
    sk = shm.SecretKey(ctx, seed=1, hamming_weight=64)
    s_ref = ref.sk_ntt(ref.sample_ternary(1, 64))
    return ctx, ref, sk, s_ref


def _operands(env, col_x, seed):
    ctx, ref, sk, s_ref = env
    rng = np.random.default_rng(seed)
    n = ctx.N // 2
    enc = shm.Encoder(ctx)
    xs = [rng.uniform(-1, 1, n) for _ in range(col_x)]
    ws = [rng.uniform(-1, 1, n) for _ in range(col_x)]
    X = [sk.encrypt_symmetric(enc.encode(x, D, 2), seed=100 + j) for j, x in enumerate(xs)]
    W = [sk.encrypt_symmetric(enc.encode(w, D, 2), seed=200 + j) for j, w in enumerate(ws)]
    return xs, ws, X, W


@pytest.mark.parametrize("num_batch", [3, 0])
def test_ctct_colpacking_bitexact(env, num_batch):
    ctx, ref, sk, s_ref = env
    col_x, row_x = 3, 5
    xs, ws, X, W = _operands(env, col_x, 7 + num_batch)
    Ql = X[0].coeff_modulus_size
    steps = sorted({(i * num_batch) % (ctx.N // 2) for i in range(1, row_x)} - {0})
    gk = sk.create_galois_keys_from_steps(steps or [1], seed=9)
    rk = sk.gen_relinkey(seed=5)
    out = shm.ct_ct_matrix_mul_colpacking(X, W, gk, rk, ctx, col_x, row_x, col_x, row_x, num_batch)
    gk_ref = {st: ref.gen_galois_key(s_ref, ref.elt(st), ref.splitmix64(9, ref.elt(st))) for st in steps}
    exp = ref.ctct_matmul_colpacking([x.numpy() for x in X], [w.numpy() for w in W], Ql, row_x, num_batch, gk_ref,
                                     ref.gen_relin_key(s_ref, 5))
    dec, enc = shm.Decryptor(ctx, sk), shm.Encoder(ctx)
    for i in range(row_x):
        assert out[i].coeff_modulus_size == Ql - 1
        assert np.array_equal(out[i].numpy(), exp[i]), f"output {i}"
        got = enc.decode(dec.decrypt(out[i]))
        want = sum(xs[j] * np.roll(ws[j], -i * num_batch) for j in range(col_x))
        assert float(np.max(np.abs(got - want))) < 1e-4, i


def test_ctct_colpacking_rejects_bad_arguments(env):
    ctx, ref, sk, s_ref = env
    xs, ws, X, W = _operands(env, 2, 3)
    gk = sk.create_galois_keys_from_steps([1], seed=9)
    rk = sk.gen_relinkey(seed=5)
    with pytest.raises(shm.capi.ShmInvalidArgument):
        shm.ct_ct_matrix_mul_colpacking(X, W, gk, rk, ctx, 2, 3, 2, 4, 1)   # row_X != row_W
    with pytest.raises(shm.capi.ShmInvalidArgument):
        shm.ct_ct_matrix_mul_colpacking(X, W, gk, rk, ctx, 2, 2, 2, 2, 1, out=[X[0], W[1]])  # aliasing


def _diag_expected(xs, ws_i, nb):
    """Slot values of the diagonal-packing product for one W ciphertext, from the
    plaintext vectors (rotate(v, s) = roll(v, -s)); test_oracle.py pins the oracle
    against the same sums."""
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


def test_ctct_diagpacking_bitexact(env):
    ctx, ref, sk, s_ref = env
    col_x, col_w, nb = 5, 3, 2
    xs, _, X, _ = _operands(env, col_x, 21)
    _, ws, _, W = _operands(env, col_w, 22)
    Ql = X[0].coeff_modulus_size
    steps = [2, 4, 6]  # babies 1..2 x nb, pre-rotation (5 - 3) nb, giant 3 nb
    gk = sk.create_galois_keys_from_steps(steps, seed=9)
    rk = sk.gen_relinkey(seed=5)
    out = shm.ct_ct_matrix_mul_diagpacking(X, W, gk, rk, ctx, col_x, col_x, col_w, col_x, nb)
    gk_ref = {st: ref.gen_galois_key(s_ref, ref.elt(st), ref.splitmix64(9, ref.elt(st))) for st in steps}
    exp = ref.ctct_matmul_diagpacking([x.numpy() for x in X], [w.numpy() for w in W], Ql, nb, gk_ref,
                                      ref.gen_relin_key(s_ref, 5))
    dec, enc = shm.Decryptor(ctx, sk), shm.Encoder(ctx)
    q_last = float(ctx.coeff_modulus[Ql - 1])
    for i in range(col_w):
        assert out[i].coeff_modulus_size == Ql - 1 and out[i].scale == D
        assert np.array_equal(out[i].numpy(), exp[i]), f"output {i}"
        got = enc.decode(dec.decrypt(out[i]))
        want = _diag_expected(xs, ws[i], nb) * D / q_last
        assert float(np.max(np.abs(got - want))) < 1e-4, i


def test_ctct_colpacking_many_digits_bitexact(oracle_lib):
    """18 key digits (one special prime, 18 data primes): the batched key switch takes
    its digits in chunks of 16 (the second chunk adds onto the first); the result is
    still the oracle's limb for limb."""
    bits = [50] + [30] * 17 + [50]
    ctx = shm.Context(12, bits, 1)
    ref = oracle_lib.Ref(12, bits, 1)
    sk = shm.SecretKey(ctx, seed=1, hamming_weight=32)
    s_ref = ref.sk_ntt(ref.sample_ternary(1, 32))
    rng = np.random.default_rng(31)
    enc = shm.Encoder(ctx)
    n = ctx.N // 2
    X = [sk.encrypt_symmetric(enc.encode(rng.uniform(-1, 1, n), 2.0 ** 30, 1), seed=300 + j) for j in range(2)]
    W = [sk.encrypt_symmetric(enc.encode(rng.uniform(-1, 1, n), 2.0 ** 30, 1), seed=310 + j) for j in range(2)]
    Ql = X[0].coeff_modulus_size
    assert Ql == 18
    gk = sk.create_galois_keys_from_steps([1, 2], seed=9)
    rk = sk.gen_relinkey(seed=5)
    out = shm.ct_ct_matrix_mul_colpacking(X, W, gk, rk, ctx, 2, 3, 2, 3, 1)
    gk_ref = {st: ref.gen_galois_key(s_ref, ref.elt(st), ref.splitmix64(9, ref.elt(st))) for st in (1, 2)}
    exp = ref.ctct_matmul_colpacking([x.numpy() for x in X], [w.numpy() for w in W], Ql, 3, 1, gk_ref,
                                     ref.gen_relin_key(s_ref, 5))
    for i in range(3):
        assert np.array_equal(out[i].numpy(), exp[i]), f"output {i}"


@pytest.fixture(scope="module")
def many(oracle_lib):
    bits = [50] + [30] * 17 + [50]
    ctx = shm.Context(12, bits, 1)
    ref = oracle_lib.Ref(12, bits, 1)
    sk = shm.SecretKey(ctx, seed=1, hamming_weight=32)
    s_ref = ref.sk_ntt(ref.sample_ternary(1, 32))
    return ctx, ref, sk, s_ref


def test_ctct_diagpacking_many_digits_bitexact(many):
    """Diagonal packing (baby-step giant-step) with 18 key digits, bit-exact."""
    ctx, ref, sk, s_ref = many
    rng = np.random.default_rng(41)
    enc = shm.Encoder(ctx)
    n = ctx.N // 2
    col_x, col_w, nb = 5, 2, 2
    X = [sk.encrypt_symmetric(enc.encode(rng.uniform(-1, 1, n), 2.0 ** 30, 1), seed=400 + j) for j in range(col_x)]
    W = [sk.encrypt_symmetric(enc.encode(rng.uniform(-1, 1, n), 2.0 ** 30, 1), seed=410 + j) for j in range(col_w)]
    Ql = X[0].coeff_modulus_size
    steps = sorted({k * nb for k in range(1, 3)} | {(col_x - 3) * nb, 3 * nb})
    gk = sk.create_galois_keys_from_steps(steps, seed=9)
    rk = sk.gen_relinkey(seed=5)
    out = shm.ct_ct_matrix_mul_diagpacking(X, W, gk, rk, ctx, col_x, col_x, col_w, col_x, nb)
    gk_ref = {st: ref.gen_galois_key(s_ref, ref.elt(st), ref.splitmix64(9, ref.elt(st))) for st in steps}
    exp = ref.ctct_matmul_diagpacking([x.numpy() for x in X], [w.numpy() for w in W], Ql, nb, gk_ref,
                                      ref.gen_relin_key(s_ref, 5))
    for i in range(col_w):
        assert np.array_equal(out[i].numpy(), exp[i]), f"output {i}"


def test_chebyshev_many_digits_batch_equals_single(many):
    """The batched products of the Chebyshev evaluation (one key switch pass for the
    group) at 18 key digits equal single calls limb for limb and decrypt to the series."""
    from numpy.polynomial import chebyshev as npc
    ctx, ref, sk, s_ref = many
    rng = np.random.default_rng(42)
    enc, dec = shm.Encoder(ctx), shm.Decryptor(ctx, sk)
    n = ctx.N // 2
    coeffs = np.array([0.1, 0.5, -0.2, 0.05, 0.02, -0.01])
    xs = [rng.uniform(-1, 1, n) for _ in range(3)]
    batch = [sk.encrypt_symmetric(enc.encode(x, 2.0 ** 30, 1), seed=500 + i) for i, x in enumerate(xs)]
    single = [sk.encrypt_symmetric(enc.encode(x, 2.0 ** 30, 1), seed=500 + i) for i, x in enumerate(xs)]
    rk = sk.gen_relinkey(seed=5)
    ev = shm.Evaluator(ctx)
    ev.evaluate_chebyshev_inplace(batch, coeffs, -1.0, 1.0, rk)
    for c in single:
        ev.evaluate_chebyshev_inplace(c, coeffs, -1.0, 1.0, rk)
    for i in range(3):
        assert np.array_equal(batch[i].numpy(), single[i].numpy()), f"member {i}"
        got = enc.decode(dec.decrypt(batch[i]))
        assert float(np.max(np.abs(got - npc.chebval(xs[i], coeffs)))) < 1e-4


def test_ctct_steps_beyond_slot_count_refused(many):
    """get_elt_from_step refuses steps of the slot count or more (galois.cu:64-69): so do
    both products, before any work."""
    ctx, ref, sk, s_ref = many
    enc = shm.Encoder(ctx)
    n = ctx.N // 2
    X = [sk.encrypt_symmetric(enc.encode(np.zeros(n), 2.0 ** 30, 1), seed=600 + j) for j in range(2)]
    gk = sk.create_galois_keys_from_steps([1], seed=9)
    rk = sk.gen_relinkey(seed=5)
    with pytest.raises(shm.capi.ShmInvalidArgument):
        shm.ct_ct_matrix_mul_colpacking(X, X, gk, rk, ctx, 2, 3, 2, 3, n // 2)  # step 2 * n/2 = n
    with pytest.raises(shm.capi.ShmInvalidArgument):
        shm.ct_ct_matrix_mul_diagpacking(X, X, gk, rk, ctx, 2, 2, 2, 2, n)
