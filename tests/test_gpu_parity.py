"""GPU parity: the HIP engine (through the C ABI) against the oracle, bit-exact.

Sizes are chosen so the oracle finishes in seconds; full-size shapes are
checked through size-independent properties in test_gpu_matmul.py.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

C1_BITS = [50] * 8 + [60]


@pytest.fixture(scope="module")
def c1(oracle_lib):
    import shmoai_amd as shm
    ctx = shm.Context(13, C1_BITS, 1)
    ref = oracle_lib.Ref(13, C1_BITS, 1)
    return ctx, ref


def _rand_limbs(ref, rng, npolys, Ql, withP=False):
    idx = ref.pidx(Ql, withP)
    out = np.empty((npolys, len(idx), ref.N), np.uint64)
    for k, pi in enumerate(idx):
        out[:, k, :] = rng.integers(0, int(ref.primes[pi]), size=(npolys, ref.N), dtype=np.uint64)
    return out.ravel()


def test_primes_match(c1):
    ctx, ref = c1
    assert np.array_equal(ctx.coeff_modulus, ref.primes)


def test_ntt_roundtrip_bitexact(c1):
    import shmoai_amd as shm
    ctx, ref = c1
    rng = np.random.default_rng(11)
    x = _rand_limbs(ref, rng, 2, 8)
    ct = shm.Ciphertext(ctx, 2, 1).load(x)
    ctx.ntt_forward(ct.data_ptr(), 2, 8, 0)
    got = ct.numpy()
    exp = ref.ntt(x, np.tile(ref.pidx(8), 2))
    assert np.array_equal(got, exp)
    ctx.ntt_inverse(ct.data_ptr(), 2, 8, 0)
    assert np.array_equal(ct.numpy(), x)


def test_ntt_rejects_misaligned_buffer(c1):
    """The tile pass reads 16-byte pairs of words: a buffer 8 bytes off that alignment is
    refused on the host, before any launch, and the words stay as they were."""
    import shmoai_amd as shm
    from shmoai_amd import capi
    ctx, ref = c1
    x = _rand_limbs(ref, np.random.default_rng(12), 2, 8)
    ct = shm.Ciphertext(ctx, 2, 1).load(x)
    with pytest.raises(capi.ShmInvalidArgument, match="16-byte aligned"):
        ctx.ntt_forward(ct.data_ptr() + 8, 1, 1, 0)
    with pytest.raises(capi.ShmInvalidArgument, match="16-byte aligned"):
        ctx.ntt_inverse(ct.data_ptr() + 8, 1, 1, 0)
    assert np.array_equal(ct.numpy(), x)


@pytest.mark.parametrize("logN,bits,n_special", [
    (12, [51, 51, 30, 20, 61], 1),
    (16, [51, 50, 46, 60], 1),
    (17, [51, 40, 51], 1),
])
def test_ntt_paths_bitexact(oracle_lib, logN, bits, n_special):
    """Primes from 20 to 61 bits and the tile splits of N = 2^12, 2^16, 2^17
    (column phase of 4 or 5 stages): the NTT must equal the oracle's bit for
    bit, including the extreme residues 0 and q - 1."""
    import shmoai_amd as shm
    ctx = shm.Context(logN, bits, n_special)
    ref = oracle_lib.Ref(logN, bits, n_special)
    N, nQ, QP = ref.N, len(bits) - n_special, len(bits)
    rng = np.random.default_rng(logN)
    for first, n in ((0, nQ), (nQ, QP - nQ)):
        x = np.empty((2, n, N), np.uint64)
        for k in range(n):
            q = int(ref.primes[first + k])
            x[0, k] = rng.integers(0, q, size=N, dtype=np.uint64)
            x[1, k] = np.where(np.arange(N) % 3 == 0, q - 1, np.where(np.arange(N) % 3 == 1, 0, q // 2))
        x = x.ravel()
        pidx = np.tile(np.arange(first, first + n), 2)
        ct = shm.Ciphertext(ctx, 2, 1)
        buf = np.zeros(2 * nQ * N, np.uint64)
        buf[:x.size] = x
        ct.load(buf)
        ctx.ntt_forward(ct.data_ptr(), 2, n, first)
        got = ct.numpy()[:x.size]
        assert np.array_equal(got, ref.ntt(x, pidx)), f"forward, primes {first}..{first + n}"
        ctx.ntt_inverse(ct.data_ptr(), 2, n, first)
        assert np.array_equal(ct.numpy()[:x.size], x), f"inverse, primes {first}..{first + n}"


def test_ntt_tile_shapes_bitexact(oracle_lib):
    """The NTT tile passes run in 512-thread blocks of radix-8 rounds by default;
    the 256-thread radix-16 form must give the same words, for the plain transforms
    (with the residues 0 and q - 1, and the oracle's forward NTT) and for a
    multiply-relinearize-rescale, whose ModDown and rescale fuse into the last pass."""
    import shmoai_amd as shm
    bits = [51] + [46] * 4 + [51] * 3 + [51, 51]
    ctx = shm.Context(16, bits, 2)
    ref = oracle_lib.Ref(16, bits, 2)
    N, nQ = ref.N, len(bits) - 2
    rng = np.random.default_rng(5)
    x = np.empty((2, nQ, N), np.uint64)
    for k in range(nQ):
        q = int(ref.primes[k])
        x[0, k] = rng.integers(0, q, size=N, dtype=np.uint64)
        x[1, k] = np.where(np.arange(N) % 2 == 0, q - 1, 0)
    x = x.ravel()
    sk = shm.SecretKey(ctx, seed=1, hamming_weight=64)
    rk = sk.gen_relinkey(seed=2)
    ev = shm.Evaluator(ctx)
    outs = {}
    assert shm.get_tuning("ntt_tile_threads") == 512
    try:
        for threads in (512, 256):
            shm.set_tuning("ntt_tile_threads", threads)
            ct = shm.Ciphertext(ctx, 2, 1).load(x)
            ctx.ntt_forward(ct.data_ptr(), 2, nQ, 0)
            fwd = ct.numpy()
            ctx.ntt_inverse(ct.data_ptr(), 2, nQ, 0)
            inv = ct.numpy()
            a = shm.Ciphertext(ctx, 2, 1).load(fwd)
            b = shm.Ciphertext(ctx, 2, 1).load(fwd.reshape(2, nQ, N)[::-1].copy())
            a.scale = b.scale = 2.0 ** 46
            ev.multiply_inplace(a, b)
            ev.relinearize_inplace(a, rk)
            ev.rescale_to_next_inplace(a)
            outs[threads] = (fwd, inv, a.numpy())
    finally:
        shm.set_tuning("ntt_tile_threads", 512)
    assert np.array_equal(outs[512][1], x)
    assert np.array_equal(outs[512][0], ref.ntt(x, np.tile(ref.pidx(nQ), 2)))
    for i, what in enumerate(("forward", "inverse", "multiply-relinearize-rescale")):
        assert np.array_equal(outs[512][i], outs[256][i]), what


def test_secret_and_keys_bitexact(c1):
    import shmoai_amd as shm
    ctx, ref = c1
    sk = shm.SecretKey(ctx, seed=1, hamming_weight=64)
    s_ref = ref.sk_ntt(ref.sample_ternary(1, 64))
    assert np.array_equal(sk.numpy(), s_ref)
    rk = sk.gen_relinkey(seed=5)
    assert np.array_equal(rk.numpy(), ref.gen_relin_key(s_ref, 5))
    gk = sk.create_galois_keys_from_steps([1, -1], seed=9)
    for step in (1, -1):
        elt = ctx.galois_elt_from_step(step)
        assert elt == ref.elt(step)
        kref = ref.gen_galois_key(s_ref, elt, ref.splitmix64(9, elt))
        assert np.array_equal(gk.numpy(elt), kref)


def test_c1_pipeline_bitexact(c1):
    """NTT -> rotate(1) -> multiply_plain -> rescale -> iNTT."""
    import shmoai_amd as shm
    ctx, ref = c1
    rng = np.random.default_rng(3)
    Ql = 8
    x = _rand_limbs(ref, rng, 2, Ql)
    sk = shm.SecretKey(ctx, seed=1, hamming_weight=64)
    s_ref = ref.sk_ntt(ref.sample_ternary(1, 64))
    gk = sk.create_galois_keys_from_steps([1], seed=9)
    elt = ref.elt(1)
    key = ref.gen_galois_key(s_ref, elt, ref.splitmix64(9, elt))
    ev = shm.Evaluator(ctx)
    ct = shm.Ciphertext(ctx, 2, 1).load(x)
    ctx.ntt_forward(ct.data_ptr(), 2, Ql, 0)
    exp = ref.ntt(x, np.tile(ref.pidx(Ql), 2))
    assert np.array_equal(ct.numpy(), exp)
    ev.rotate_vector_inplace(ct, 1, gk)
    exp = ref.rotate(exp, Ql, elt, key)
    assert np.array_equal(ct.numpy(), exp)
    c = float(np.random.default_rng(4).uniform(-1, 1))
    pt = shm.Encoder(ctx).encode(c, 2.0 ** 50, 1)
    pt_ref = ref.encode_scalar(c, Ql, 2.0 ** 50)
    assert np.array_equal(pt.numpy(), pt_ref)
    ev.multiply_plain_inplace(ct, pt)
    exp = ref.multiply_plain(exp, pt_ref, Ql)
    assert np.array_equal(ct.numpy(), exp)
    ev.rescale_to_next_inplace(ct)
    exp = ref.rescale(exp, Ql)
    assert np.array_equal(ct.numpy(), exp)
    ctx.ntt_inverse(ct.data_ptr(), 2, Ql - 1, 0)
    exp = ref.intt(exp, np.tile(ref.pidx(Ql - 1), 2))
    assert np.array_equal(ct.numpy(), exp)


def test_encrypt_decrypt_encode_bitexact(c1):
    import shmoai_amd as shm
    ctx, ref = c1
    rng = np.random.default_rng(21)
    v = rng.uniform(-1, 1, ctx.N // 2)
    enc = shm.Encoder(ctx)
    pt = enc.encode(v, 2.0 ** 40, 1)
    pt_ref = ref.encode(v, 8, 2.0 ** 40)
    assert np.array_equal(pt.numpy(), pt_ref)
    sk = shm.SecretKey(ctx, seed=1, hamming_weight=64)
    s_ref = ref.sk_ntt(ref.sample_ternary(1, 64))
    ct = sk.encrypt_symmetric(pt, seed=77)
    ct_ref = ref.encrypt_symmetric(s_ref, pt_ref, 8, 77)
    assert np.array_equal(ct.numpy(), ct_ref)
    dec = shm.Decryptor(ctx, sk).decrypt(ct)
    dec_ref = ref.decrypt(s_ref, ct_ref, 2, 8)
    assert np.array_equal(dec.numpy(), dec_ref)
    out = enc.decode(dec)
    re_ref, _ = ref.decode(dec_ref, 8, 2.0 ** 40)
    assert np.max(np.abs(out - re_ref)) <= 1e-12
    assert np.max(np.abs(out - v)) < 1e-6


@pytest.mark.parametrize("bits,n_special,Ql", [(C1_BITS, 1, 8), ([50] * 8 + [60] * 3, 3, 6),
                                               ([51] + [46] * 8 + [51] * 3, 3, 7)])
def test_hoisting_bitexact(oracle_lib, bits, n_special, Ql):
    """evaluate.cu:1727-1920 hoisting_inplace: sum of rotations, one ModUp of c1,
    permuted digits per step, summed inner products, one ModDown. The third case has
    every prime below 2^52 (the reference layer's 46/51-bit chain), where the base
    conversions take 26-bit split products, and a last digit of one prime (7 = 3 + 3 + 1)."""
    import shmoai_amd as shm
    ctx = shm.Context(13, bits, n_special)
    ref = oracle_lib.Ref(13, bits, n_special)
    rng = np.random.default_rng(8)
    N, steps = ref.N, [1, 2, -3]
    x = _rand_limbs(ref, rng, 2, Ql)
    sk = shm.SecretKey(ctx, seed=1, hamming_weight=64)
    s_ref = ref.sk_ntt(ref.sample_ternary(1, 64))
    gk = sk.create_galois_keys_from_steps(steps, seed=9)
    ct = shm.Ciphertext(ctx, 2, ctx.chain(Ql)).load(x)
    shm.Evaluator(ctx).hoisting_inplace(ct, gk, steps)
    # oracle composition
    QlP = Ql + ref.nP
    qs = np.array([int(ref.primes[k]) for k in ref.pidx(Ql, True)], dtype=np.uint64)[:, None]
    qQ = qs[:Ql]
    c0, c1 = x[:Ql * N], x[Ql * N:]
    d = ref.modup(c1, Ql)
    acc = np.zeros((2, QlP, N), np.uint64)
    c0sum = np.zeros((Ql, N), np.uint64)
    for st in steps:
        elt = ref.elt(st)
        key = ref.gen_galois_key(s_ref, elt, ref.splitmix64(9, elt))
        dp = ref.apply_galois(d, elt, ref.beta(Ql) * QlP)
        cx = ref.ks_inner(dp, key, Ql).reshape(2, QlP, N)
        acc = (acc + cx) % qs
        c0sum = (c0sum + ref.apply_galois(c0, elt, Ql).reshape(Ql, N)) % qQ
    k0 = ref.moddown(acc[0].ravel(), Ql).reshape(Ql, N)
    k1 = ref.moddown(acc[1].ravel(), Ql)
    exp = np.concatenate([((c0sum + k0) % qQ).ravel(), k1])
    assert np.array_equal(ct.numpy(), exp)


def test_keyswitch_refuses_missing_key_digits():
    """Keys hold nQ / nP digits (secretkey.cu:328-373); with nQ = 8, nP = 3 that is 2, and
    a key switch at Ql = 7 needs 3: hoisting, rotation and relinearisation refuse it
    before any kernel is launched."""
    import shmoai_amd as shm
    ctx = shm.Context(13, [51] + [46] * 7 + [51] * 3, 3)
    sk = shm.SecretKey(ctx, seed=1, hamming_weight=64)
    gk = sk.create_galois_keys_from_steps([1], seed=9)
    rk = sk.gen_relinkey(seed=10)
    ev = shm.Evaluator(ctx)
    ct = shm.Ciphertext(ctx, 2, ctx.chain(7))
    with pytest.raises(shm.capi.ShmLogicError):
        ev.hoisting_inplace(ct, gk, [1])
    with pytest.raises(shm.capi.ShmLogicError):
        ev.rotate_vector_inplace(ct, 1, gk)
    ct3 = shm.Ciphertext(ctx, 3, ctx.chain(7))
    with pytest.raises(shm.capi.ShmLogicError):
        ev.relinearize_inplace(ct3, rk)
    # one level lower (Ql = 6, 2 digits) is fine
    ok = shm.Ciphertext(ctx, 2, ctx.chain(6))
    ev.rotate_vector_inplace(ok, 1, gk)
    shm.synchronize()
