"""GPU bootstrapping: bit-exact pieces against the oracle, tolerance on the
whole pipeline.

- modraise (Bootstrapper.cu:3287-3356) and the X^(N/2) multiply are integer
  maps: bit-exact against the oracle's coefficient-domain restatement;
- slot-to-coefficient: decrypted coefficients = F * scale * (R z) split into
  real and imaginary halves (closed form of the transform);
- EvalMod: sin(2 pi x) near the integers;
- bootstrap_s2c_first end to end: max |decrypt(bootstrap(ct)) - z| <= 2e-5
  over all slots (the pass bar of the bootstrap row).
"""
import numpy as np
import pytest

import shmoai_amd as shm

pytestmark = pytest.mark.gpu

# N = 2^13: q0 51 bits, 3 x 46-bit S2C primes, 14 x 51-bit primes above
SMALL_BITS = [51] + [46] * 3 + [51] * 14 + [51] * 3
SMALL_NP = 3


def bitrev_perm(n):
    b = n.bit_length() - 1
    return np.array([int(format(i, f"0{b}b")[::-1], 2) for i in range(n)])


@pytest.fixture(scope="module")
def small(oracle_lib):
    ctx = shm.Context(13, SMALL_BITS, SMALL_NP)
    ref = oracle_lib.Ref(13, SMALL_BITS, SMALL_NP)
    sk = shm.SecretKey(ctx, seed=21, hamming_weight=64)
    boot = shm.Bootstrapper(ctx)
    gk = sk.create_galois_keys_from_steps(boot.rotation_steps(), seed=22)
    rk = sk.gen_relinkey(seed=23)
    return ctx, ref, sk, boot, gk, rk


def test_modraise_bitexact(small):
    ctx, ref, _, boot, *_ = small
    rng = np.random.default_rng(31)
    q0 = int(ctx.coeff_modulus[0])
    low = rng.integers(0, q0, size=2 * ctx.N, dtype=np.uint64)
    ct = shm.Ciphertext(ctx, 2, ctx.size_q).load(low)
    boot.modraise_inplace(ct)
    assert ct.chain_index == 1
    assert np.array_equal(ct.numpy(), ref.modraise(low))


@pytest.mark.parametrize("negate", [False, True])
def test_multiply_by_i_bitexact(small, negate):
    ctx, ref, *_ = small
    rng = np.random.default_rng(32)
    Ql = 5
    a = np.empty((2, Ql, ctx.N), np.uint64)
    for k in range(Ql):
        a[:, k, :] = rng.integers(0, int(ctx.coeff_modulus[k]), size=(2, ctx.N), dtype=np.uint64)
    a = a.ravel()
    ct = shm.Ciphertext(ctx, 2, ctx.chain(Ql)).load(a)
    shm.Evaluator(ctx).multiply_by_i_inplace(ct, negate=negate)
    power = (3 if negate else 1) * ctx.N // 2
    assert np.array_equal(ct.numpy(), ref.mul_monomial(a, Ql, 2, power))


def _encrypt(ctx, sk, z, scale, Ql, seed):
    enc = shm.Encoder(ctx)
    return sk.encrypt_symmetric(enc.encode(z, scale, ctx.chain(Ql)), seed=seed)


def test_slot_to_coeff_closed_form(small):
    ctx, ref, sk, boot, gk, _ = small
    n = ctx.N // 2
    rng = np.random.default_rng(33)
    z = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    scale = 2.0 ** 46
    ct = _encrypt(ctx, sk, z, scale, 4, 34)
    boot.slot_to_coeff(ct, gk)
    assert ct.coeff_modulus_size == 1 and ct.scale == scale
    pt = shm.Decryptor(ctx, sk).decrypt(ct).numpy()
    q0 = int(ctx.coeff_modulus[0])
    c = ref.intt(pt[:ctx.N].copy(), [0]).astype(object)
    c = np.array([int(v) - q0 if int(v) > q0 // 2 else int(v) for v in c], dtype=np.float64)
    v = z[bitrev_perm(n)] / 16.0
    err = np.max(np.abs(c / scale - np.concatenate([v.real, v.imag])))
    assert err < 1e-6, err


def test_eval_mod_sine(small):
    ctx, ref, sk, boot, gk, rk = small
    n = ctx.N // 2
    K = boot.K
    rng = np.random.default_rng(35)
    ints = rng.integers(-20, 21, size=n)
    eps = rng.uniform(-1 / 512, 1 / 512, n)
    x = ints + eps
    q0 = float(ctx.coeff_modulus[0])
    ct = _encrypt(ctx, sk, x, q0, ctx.size_q - 3, 36)
    boot.eval_mod(ct, rk)
    assert ct.scale == q0 and ct.coeff_modulus_size == ctx.size_q - 13
    got = shm.Encoder(ctx).decode(shm.Decryptor(ctx, sk).decrypt(ct))
    err = np.max(np.abs(got - np.sin(2 * np.pi * x)))
    assert err < 1e-6, err


def test_eval_mod_batch_bitexact(small):
    """Batched EvalMod (one key switch and rescale pass per step for the group)
    equals EvalMod applied to each ciphertext alone, limb for limb."""
    ctx, ref, sk, boot, gk, rk = small
    n = ctx.N // 2
    rng = np.random.default_rng(38)
    q0 = float(ctx.coeff_modulus[0])
    xs = [rng.integers(-20, 21, size=n) + rng.uniform(-1 / 512, 1 / 512, n) for _ in range(3)]
    batch = [_encrypt(ctx, sk, x, q0, ctx.size_q - 3, 60 + i) for i, x in enumerate(xs)]
    single = [_encrypt(ctx, sk, x, q0, ctx.size_q - 3, 60 + i) for i, x in enumerate(xs)]
    boot.eval_mod_batch(batch, rk)
    for b, t, x in zip(batch, single, xs):
        boot.eval_mod(t, rk)
        assert b.coeff_modulus_size == t.coeff_modulus_size and b.scale == t.scale
        assert np.array_equal(b.numpy(), t.numpy())
        got = shm.Encoder(ctx).decode(shm.Decryptor(ctx, sk).decrypt(b))
        assert np.max(np.abs(got - np.sin(2 * np.pi * x))) < 1e-6
    with pytest.raises(shm.capi.ShmInvalidArgument):
        boot.eval_mod_batch([batch[0], _encrypt(ctx, sk, xs[0], q0, ctx.size_q - 4, 70)], rk)


def test_bootstrap_s2c_first_small(small):
    ctx, ref, sk, boot, gk, rk = small
    n = ctx.N // 2
    rng = np.random.default_rng(37)
    zs = [rng.uniform(-1, 1, n), rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)]
    scale = 2.0 ** 46
    cts = [_encrypt(ctx, sk, z, scale, 4, 40 + i) for i, z in enumerate(zs)]
    out = boot.bootstrap_s2c_first(cts, gk, rk)
    dec, enc = shm.Decryptor(ctx, sk), shm.Encoder(ctx)
    for z, o in zip(zs, out):
        assert o.scale == scale
        assert o.coeff_modulus_size == ctx.size_q - 13
        got = enc.decode(dec.decrypt(o), complex_out=True)
        err = np.max(np.abs(got - z))
        print(f"bootstrap N=2^13 max err {err:.3e}", flush=True)
        assert err <= 2e-5, err
    # inputs unchanged, wrong level rejected
    assert cts[0].coeff_modulus_size == 4
    bad = _encrypt(ctx, sk, zs[0], scale, 5, 50)
    with pytest.raises(shm.capi.ShmInvalidArgument):
        boot.bootstrap_s2c_first([bad], gk, rk)


def test_bootstrap_group_equals_single(small):
    """Bootstrapping three ciphertexts in one call (one group: batched linear
    transforms and EvalMod) gives exactly the limbs of bootstrapping each alone."""
    ctx, ref, sk, boot, gk, rk = small
    n = ctx.N // 2
    rng = np.random.default_rng(39)
    zs = [rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n) for _ in range(3)]
    cts = [_encrypt(ctx, sk, z, 2.0 ** 46, 4, 80 + i) for i, z in enumerate(zs)]
    group = boot.bootstrap_s2c_first(cts, gk, rk)
    for ct, g in zip(cts, group):
        alone = boot.bootstrap_s2c_first([ct], gk, rk)[0]
        assert g.coeff_modulus_size == alone.coeff_modulus_size and g.scale == alone.scale
        assert np.array_equal(g.numpy(), alone.numpy())


def test_bootstrap_s2c_first_reference_params():
    """N = 2^16, Q = 1 x 51 + 20 x 46 + 14 x 51 bits, P = 5 x 51 bits (dnum 7),
    Hamming weight 192, scale 2^46, inputs with 4 primes
    (test_shared_a.cuh:1207-1274 s2c_first_bootstrap_test parameters)."""
    bits = [51] + [46] * 20 + [51] * 14 + [51] * 5
    ctx = shm.Context(16, bits, 5)
    sk = shm.SecretKey(ctx, seed=61, hamming_weight=192)
    boot = shm.Bootstrapper(ctx)
    gk = sk.create_galois_keys_from_steps(boot.rotation_steps(), seed=62)
    rk = sk.gen_relinkey(seed=63)
    n = ctx.N // 2
    rng = np.random.default_rng(64)
    zs = [rng.uniform(-1, 1, n), rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)]
    cts = [_encrypt(ctx, sk, z, 2.0 ** 46, 4, 70 + i) for i, z in enumerate(zs)]
    out = boot.bootstrap_s2c_first(cts, gk, rk)
    dec, enc = shm.Decryptor(ctx, sk), shm.Encoder(ctx)
    for z, o in zip(zs, out):
        assert o.coeff_modulus_size == 22 and o.scale == 2.0 ** 46
        err = np.max(np.abs(enc.decode(dec.decrypt(o), complex_out=True) - z))
        print(f"bootstrap N=2^16 max err {err:.3e}", flush=True)
        assert err <= 2e-5, err
