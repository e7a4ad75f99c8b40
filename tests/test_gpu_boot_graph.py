"""Batched bootstrapping over many chunks at the reference parameters (N = 2^16,
Q = 1 x 51 + 20 x 46 + 14 x 51 bits, P = 5 x 51 bits, HW 192, inputs with 4 primes):

- the chunked shared-a entry point (k chunks of n = 8 in one call) equals one call per
  chunk, limb for limb;
- one full pass of 768 ciphertexts (96 chunks of n = 8, the batch of one BERT-base layer,
  test_shared_a.cuh:1336-1392) decrypts within 2e-5 of the inputs on a sample of outputs
  spread over the chunks;
- graph-captured chunks are switched off (their first replay faulted; creation refuses).
"""
import numpy as np
import pytest

import shmoai_amd as shm

pytestmark = pytest.mark.gpu

BITS = [51] + [46] * 20 + [51] * 14 + [51] * 5
SCALE = 2.0 ** 46
GRAPH_OFF = ("graph-captured bootstrapping is switched off: its first replay on an MI355X ended in an "
             "illegal memory access whose cause is not established")


@pytest.fixture(scope="module")
def c3():
    ctx = shm.Context(16, BITS, 5)
    sk = shm.SecretKey(ctx, seed=1, hamming_weight=192)
    boot = shm.Bootstrapper(ctx)
    steps = boot.rotation_steps()
    rk = sk.gen_relinkey(seed=3)
    boot.prepare()
    keys = shm.SharedAKeys(ctx, sk, 8, 192, 12, ctx.chain(4))
    keys.add_galois(steps, seed=13)
    return ctx, sk, boot, steps, rk, keys


def _inputs(ctx, sk, count, seed):
    enc = shm.Encoder(ctx)
    rng = np.random.default_rng(seed)
    zs = [rng.uniform(-1, 1, ctx.N // 2) + 1j * rng.uniform(-1, 1, ctx.N // 2) for _ in range(count)]
    cts = [sk.encrypt_symmetric(enc.encode(z, SCALE, ctx.chain(4)), seed=1000 + seed * 1000 + i)
           for i, z in enumerate(zs)]
    return zs, cts


def test_shared_a_chunked_equals_single_calls(c3):
    ctx, sk, boot, steps, rk, keys = c3
    zs, cts = _inputs(ctx, sk, 16, 1)
    single = []
    for c in range(2):
        single += shm.bootstrap_shared_a_s2c_first(boot, keys, cts[8 * c:8 * (c + 1)], rk)
    chunked = shm.bootstrap_shared_a_s2c_first_chunks(boot, keys, cts, rk)
    for i in range(16):
        assert np.array_equal(chunked[i].numpy(), single[i].numpy()), f"member {i}"
    dec, enc = shm.Decryptor(ctx, sk), shm.Encoder(ctx)
    err = max(float(np.max(np.abs(enc.decode(dec.decrypt(chunked[i]), complex_out=True) - zs[i]))) for i in (0, 15))
    assert err < 2e-5, err
    with pytest.raises(shm.capi.ShmInvalidArgument):
        shm.bootstrap_shared_a_s2c_first_chunks(boot, keys, cts[:7], rk)


def test_full_768_shared_a_pass(c3):
    ctx, sk, boot, steps, rk, keys = c3
    zs, cts = _inputs(ctx, sk, 768, 3)
    outs = shm.bootstrap_shared_a_s2c_first_chunks(boot, keys, cts, rk)
    dec, enc = shm.Decryptor(ctx, sk), shm.Encoder(ctx)
    sample = [0, 7, 200, 383, 384, 571, 760, 767]
    err = max(float(np.max(np.abs(enc.decode(dec.decrypt(outs[i]), complex_out=True) - zs[i]))) for i in sample)
    assert err < 2e-5, err
    assert all(o.chain_index == outs[0].chain_index for o in outs)


def test_shared_a_n16_chunk(c3):
    """A shared-a chunk of n = 16 (17 Sh-Rot outputs, 32 EvalMod members) equals the
    chunked entry over the same 16 and decrypts within 2e-5."""
    ctx, sk, boot, steps, rk, keys = c3
    keys16 = shm.SharedAKeys(ctx, sk, 16, 192, 56, ctx.chain(4))
    keys16.add_galois(steps, seed=17)
    zs, cts = _inputs(ctx, sk, 16, 5)
    single = shm.bootstrap_shared_a_s2c_first(boot, keys16, cts, rk)
    chunked = shm.bootstrap_shared_a_s2c_first_chunks(boot, keys16, cts, rk)
    for i in (0, 9, 15):
        assert np.array_equal(chunked[i].numpy(), single[i].numpy()), f"member {i}"
    dec, enc = shm.Decryptor(ctx, sk), shm.Encoder(ctx)
    err = max(float(np.max(np.abs(enc.decode(dec.decrypt(single[i]), complex_out=True) - zs[i]))) for i in range(16))
    assert err < 2e-5, err


def test_graph_capture_is_refused(c3):
    ctx, sk, boot, steps, rk, keys = c3
    with pytest.raises(shm.capi.ShmLogicError):
        shm.BootGraph.shared_a(boot, keys, rk, SCALE)


@pytest.mark.skip(reason=GRAPH_OFF)
def test_shared_a_graph_equals_eager(c3):
    ctx, sk, boot, steps, rk, keys = c3
    zs, cts = _inputs(ctx, sk, 16, 1)
    eager = shm.bootstrap_shared_a_s2c_first_chunks(boot, keys, cts, rk)
    g = shm.BootGraph.shared_a(boot, keys, rk, SCALE)
    outs = g.run(cts)
    for i in range(16):
        assert np.array_equal(outs[i].numpy(), eager[i].numpy()), f"graph {i}"


@pytest.mark.skip(reason=GRAPH_OFF)
def test_ordinary_graph_equals_eager(c3):
    ctx, sk, boot, steps, rk, keys = c3
    gk = sk.create_galois_keys_from_steps(steps, seed=2)
    zs, cts = _inputs(ctx, sk, 4, 2)
    eager = boot.bootstrap_s2c_first(cts, gk, rk)
    outs = shm.BootGraph.ordinary(boot, gk, rk, 2, SCALE).run(cts)
    for i in range(4):
        assert np.array_equal(outs[i].numpy(), eager[i].numpy()), f"member {i}"
