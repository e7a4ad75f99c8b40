"""GPU shared-a format: bit-exact against the oracle, and the batched bootstrap.

Bit-exact (same injected seeds / same inputs):
- auxiliary secrets, the format key (FmtSwitch.cu:164-270), a shared Galois key
  (ShRotate.cu:54-162) and the backward keys (gen_kswitch_key s'_i -> s);
- fmt_switch_forward against the oracle restatement;
- fmt_switch_backward member i against keyswitch(a, backward key i) + (b_i, 0);
- Sh-Rot member i against an ordinary rotation of (b_i, a) with the key
  (B_{g,i}, A_g) (ShRotate.cu:201-257 applies exactly that key to every member);
- batch add / multiply_plain / rescale / modraise against per-member oracle ops.
Tolerance: bootstrap_shared_a_s2c_first, max |decrypt - z| <= 2e-5.
"""
import numpy as np
import pytest

import shmoai_amd as shm

pytestmark = pytest.mark.gpu

BITS = [51] + [46] * 3 + [51] * 16 + [51] * 4
NP = 4
N_MEMBERS = 3
SEED = 77
HW = 64


def splitmix64(key, ctr):
    M = (1 << 64) - 1
    z = (key + (ctr + 1) * 0x9E3779B97F4A7C15) & M
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


@pytest.fixture(scope="module")
def env(oracle_lib):
    ctx = shm.Context(13, BITS, NP)
    ref = oracle_lib.Ref(13, BITS, NP)
    sk = shm.SecretKey(ctx, seed=5, hamming_weight=HW)
    keys = shm.SharedAKeys(ctx, sk, N_MEMBERS, HW, SEED, ctx.chain(4))
    return ctx, ref, sk, keys


def _rand(ref, rng, npolys, Ql):
    a = np.empty((npolys, Ql, ref.N), np.uint64)
    for k in range(Ql):
        a[:, k, :] = rng.integers(0, int(ref.primes[k]), size=(npolys, ref.N), dtype=np.uint64)
    return a


def test_shared_keys_bitexact(env):
    ctx, ref, sk, keys = env
    s_ref = ref.sk_ntt(ref.sample_ternary(5, HW))
    assert np.array_equal(sk.numpy(), s_ref)
    aux = []
    for i in range(N_MEMBERS):
        coeffs = ref.sample_ternary(splitmix64(SEED, i), HW)
        assert np.array_equal(keys.aux_coeffs(i).astype(np.int64), coeffs)
        aux.append(ref.sk_ntt(coeffs))
    a_fmt, B = keys.export_format()
    ra, rB = ref.gen_fmt_key(s_ref, aux, 4, SEED)
    assert np.array_equal(a_fmt, ra) and np.array_equal(B, rB)
    keys.add_galois([3], seed=91)
    elt = ctx.galois_elt_from_step(3)
    A, GB = keys.export_galois(elt)
    rA, rGB = ref.gen_shared_galois_key(aux, elt, 91)
    assert np.array_equal(A, rA) and np.array_equal(GB, rGB)


def test_fmt_switch_forward_bitexact(env):
    ctx, ref, sk, keys = env
    rng = np.random.default_rng(1)
    Ql = 4
    cts = [_rand(ref, rng, 2, Ql).ravel() for _ in range(N_MEMBERS)]
    hs = [shm.Ciphertext(ctx, 2, ctx.chain(Ql)).load(c) for c in cts]
    for h in hs:
        h.scale = 2.0 ** 46
    batch = keys.fmt_switch_forward(hs)
    assert batch.chain_index == ctx.chain(Ql) and batch.scale == 2.0 ** 46
    a_fmt, B = keys.export_format()
    assert np.array_equal(batch.numpy(), ref.fmt_switch_forward(cts, Ql, a_fmt, B))


def test_fmt_switch_backward_bitexact(env):
    ctx, ref, sk, keys = env
    rng = np.random.default_rng(2)
    Ql = 6
    data = _rand(ref, rng, N_MEMBERS + 1, Ql)
    batch = shm.SharedABatch(ctx, N_MEMBERS, ctx.chain(Ql)).load(data.ravel())
    outs = keys.fmt_switch_backward(batch)
    s_ref = ref.sk_ntt(ref.sample_ternary(5, HW))
    for i in range(N_MEMBERS):
        aux_i = ref.sk_ntt(ref.sample_ternary(splitmix64(SEED, i), HW))
        key = ref.gen_kswitch_key(s_ref, aux_i, splitmix64(SEED, 2 * N_MEMBERS + 2 + i))
        k0, k1 = ref.keyswitch(data[N_MEMBERS].ravel(), key, Ql)
        c0 = ref.add(data[i].ravel(), k0, Ql, 1)
        assert np.array_equal(outs[i].numpy(), np.concatenate([c0, k1])), f"member {i}"


def test_shared_rotate_bitexact(env):
    ctx, ref, sk, keys = env
    rng = np.random.default_rng(3)
    Ql = 5
    keys.add_galois([3], seed=91)
    elt = ctx.galois_elt_from_step(3)
    A, GB = keys.export_galois(elt)
    N, QP, dnum, n = ctx.N, ctx.size_q + ctx.size_p, ctx.dnum, N_MEMBERS
    data = _rand(ref, rng, n + 1, Ql)
    batch = shm.SharedABatch(ctx, n, ctx.chain(Ql)).load(data.ravel())
    keys.rotate_inplace(batch, 3)
    got = batch.numpy().reshape(n + 1, Ql * N)
    A = A.reshape(dnum, QP * N)
    GB = GB.reshape(dnum, n, QP * N)
    for i in range(n):
        key = np.stack([np.concatenate([GB[g, i], A[g]]) for g in range(dnum)]).ravel()
        ct = np.concatenate([data[i].ravel(), data[n].ravel()])
        exp = ref.rotate(ct, Ql, elt, key).reshape(2, Ql * N)
        assert np.array_equal(got[i], exp[0]), f"b member {i}"
        assert np.array_equal(got[n], exp[1])


def test_batch_ops_bitexact(env):
    ctx, ref, sk, keys = env
    rng = np.random.default_rng(4)
    Ql, n, N = 5, N_MEMBERS, ctx.N
    x = _rand(ref, rng, n + 1, Ql)
    y = _rand(ref, rng, n + 1, Ql)
    pt = _rand(ref, rng, 1, Ql)
    bx = shm.SharedABatch(ctx, n, ctx.chain(Ql)).load(x.ravel())
    by = shm.SharedABatch(ctx, n, ctx.chain(Ql)).load(y.ravel())
    bx.add_inplace(by)
    exp = ref.add(x.ravel(), y.ravel(), Ql, n + 1)
    assert np.array_equal(bx.numpy(), exp)
    p = shm.Plaintext(ctx, ctx.chain(Ql)).load(pt.ravel())
    bx.multiply_plain_inplace(p)
    exp = ref.multiply_plain(exp, pt.ravel(), Ql, n + 1)
    assert np.array_equal(bx.numpy(), exp)
    bx.rescale_to_next_inplace()
    exp = ref.rescale(exp, Ql, n + 1)
    assert np.array_equal(bx.numpy(), exp)
    low = _rand(ref, rng, n + 1, 1)
    bl = shm.SharedABatch(ctx, n, ctx.size_q).load(low.ravel())
    bl.modraise_inplace()
    got = bl.numpy().reshape(n + 1, ctx.size_q * N)
    for i in range(0, n + 1, 2):
        pair = np.concatenate([low[i].ravel(), low[(i + 1) % (n + 1)].ravel()])
        exp2 = ref.modraise(pair).reshape(2, -1)
        assert np.array_equal(got[i], exp2[0])


def _encrypt(ctx, sk, z, Ql, seed):
    return sk.encrypt_symmetric(shm.Encoder(ctx).encode(z, 2.0 ** 46, ctx.chain(Ql)), seed=seed)


def test_bootstrap_shared_a_small(env):
    ctx, ref, sk, keys = env
    boot = shm.Bootstrapper(ctx)
    keys.add_galois(boot.rotation_steps(), seed=92)
    rk = sk.gen_relinkey(seed=93)
    n = ctx.N // 2
    rng = np.random.default_rng(6)
    zs = [rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n) for _ in range(N_MEMBERS)]
    cts = [_encrypt(ctx, sk, z, 4, 60 + i) for i, z in enumerate(zs)]
    out = shm.bootstrap_shared_a_s2c_first(boot, keys, cts, rk)
    dec, enc = shm.Decryptor(ctx, sk), shm.Encoder(ctx)
    for z, o in zip(zs, out):
        assert o.scale == 2.0 ** 46 and o.coeff_modulus_size == ctx.size_q - 13
        err = np.max(np.abs(enc.decode(dec.decrypt(o), complex_out=True) - z))
        print(f"shared-a bootstrap N=2^13 max err {err:.3e}", flush=True)
        assert err <= 2e-5, err


def test_bootstrap_shared_a_reference_params():
    """N = 2^16, Q = 1 x 51 + 20 x 46 + 14 x 51, P = 5 x 51, HW 192, n = 4
    (test_shared_a.cuh:1207-1274)."""
    bits = [51] + [46] * 20 + [51] * 14 + [51] * 5
    ctx = shm.Context(16, bits, 5)
    sk = shm.SecretKey(ctx, seed=61, hamming_weight=192)
    boot = shm.Bootstrapper(ctx)
    keys = shm.SharedAKeys(ctx, sk, 4, 192, 65, ctx.chain(4))
    keys.add_galois(boot.rotation_steps(), seed=66)
    rk = sk.gen_relinkey(seed=63)
    n = ctx.N // 2
    rng = np.random.default_rng(64)
    zs = [rng.uniform(-1, 1, n) for _ in range(4)]
    cts = [_encrypt(ctx, sk, z, 4, 70 + i) for i, z in enumerate(zs)]
    out = shm.bootstrap_shared_a_s2c_first(boot, keys, cts, rk)
    dec, enc = shm.Decryptor(ctx, sk), shm.Encoder(ctx)
    for z, o in zip(zs, out):
        assert o.coeff_modulus_size == 22 and o.scale == 2.0 ** 46
        err = np.max(np.abs(enc.decode(dec.decrypt(o), complex_out=True) - z))
        print(f"shared-a bootstrap N=2^16 max err {err:.3e}", flush=True)
        assert err <= 2e-5, err
