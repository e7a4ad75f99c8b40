"""Bootstrapping server with continuous batching (no counterpart in the reference).

- requests of uneven sizes (and one at another scale), submitted from two threads
  while others run, each get exactly the limbs of bootstrapping their ciphertexts
  alone, and decrypt within 2e-5;
- admission by memory: with a budget below one request, requests enter one at a
  time (so groups never mix requests) and the results are unchanged;
- a request with the wrong level is refused at submit and the server keeps serving;
- two workers (two streams, groups running concurrently) give the same limbs, and
  admission by the device's free memory alone (no budget) admits everything;
- packed requests of uneven slot counts share ciphertexts: each decrypts, at its slot
  offset, within 2e-5 of its input and of the same request bootstrapped alone, with
  zeros in the rest of its block; requests without the Galois keys the packing needs
  are refused.
"""
import threading
import time

import numpy as np
import pytest

import shmoai_amd as shm

pytestmark = pytest.mark.gpu

# N = 2^13: q0 51 bits, 3 x 46-bit S2C primes, 14 x 51-bit primes above
SMALL_BITS = [51] + [46] * 3 + [51] * 14 + [51] * 3
SMALL_NP = 3


@pytest.fixture(scope="module")
def env():
    ctx = shm.Context(13, SMALL_BITS, SMALL_NP)
    sk = shm.SecretKey(ctx, seed=21, hamming_weight=64)
    boot = shm.Bootstrapper(ctx)
    gk = sk.create_galois_keys_from_steps(boot.rotation_steps(), seed=22)
    rk = sk.gen_relinkey(seed=23)
    return ctx, sk, boot, gk, rk


def _requests(env, sizes, seed, scales=None):
    ctx, sk, *_ = env
    enc = shm.Encoder(ctx)
    rng = np.random.default_rng(seed)
    n = ctx.N // 2
    reqs, k = [], 0
    for r, m in enumerate(sizes):
        sc = scales[r] if scales else 2.0 ** 46
        zs = [rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n) for _ in range(m)]
        cts = [sk.encrypt_symmetric(enc.encode(z, sc, ctx.chain(4)), seed=500 + seed * 100 + (k := k + 1))
               for z in zs]
        reqs.append((zs, cts))
    return reqs


def _check_alone(env, reqs, results, tol_check=True):
    ctx, sk, boot, gk, rk = env
    dec, enc = shm.Decryptor(ctx, sk), shm.Encoder(ctx)
    for (zs, cts), outs in zip(reqs, results):
        assert len(outs) == len(cts)
        for z, ct, o in zip(zs, cts, outs):
            alone = boot.bootstrap_s2c_first([ct], gk, rk)[0]
            assert o.scale == alone.scale and o.coeff_modulus_size == alone.coeff_modulus_size
            assert np.array_equal(o.numpy(), alone.numpy())
        if tol_check:
            got = enc.decode(dec.decrypt(outs[0]), complex_out=True)
            assert float(np.max(np.abs(got - zs[0]))) <= 2e-5


def test_server_mix_equals_alone(env):
    ctx, sk, boot, gk, rk = env
    sizes = [3, 1, 5, 2, 2]
    scales = [2.0 ** 46] * 4 + [2.0 ** 45]
    reqs = _requests(env, sizes, 1, scales)
    srv = shm.BootstrapServer(boot, gk, rk, mem_budget=64 << 30, max_group=4)
    ids = [None] * len(reqs)

    def late():
        time.sleep(0.05)
        for r in (1, 3, 4):
            ids[r] = srv.submit(reqs[r][1])
            time.sleep(0.01)

    t = threading.Thread(target=late)
    t.start()
    for r in (0, 2):
        ids[r] = srv.submit(reqs[r][1])
    t.join()
    results, lats = [], []
    for rid in ids:
        outs, lat = srv.wait(rid)
        results.append(outs)
        lats.append(lat)
    st = srv.stats()
    print("server stats", st, "latencies ms", [round(x, 1) for x in lats], flush=True)
    assert st["submitted"] == 5 and st["completed"] == 5 and st["failed"] == 0
    assert st["ciphertexts"] == sum(sizes)
    assert st["in_flight_bytes"] == 0
    assert 0 < st["p50_ms"] <= st["p99_ms"]
    assert st["groups"] >= 4  # 11 slots at scale 2^46 in groups of <= 4, 2 at 2^45
    _check_alone(env, reqs, results)
    srv.close()


def test_server_memory_admission(env):
    ctx, sk, boot, gk, rk = env
    sizes = [2, 3]
    reqs = _requests(env, sizes, 2)
    srv = shm.BootstrapServer(boot, gk, rk, mem_budget=1, max_group=8)
    ids = [srv.submit(cts) for _, cts in reqs]
    results = [srv.wait(i)[0] for i in ids]
    st = srv.stats()
    assert st["groups"] == 2, st  # one request admitted at a time: no mixed group
    _check_alone(env, reqs, results, tol_check=False)
    srv.close()


def test_server_rejects_wrong_level(env):
    ctx, sk, boot, gk, rk = env
    enc = shm.Encoder(ctx)
    n = ctx.N // 2
    bad = sk.encrypt_symmetric(enc.encode(np.zeros(n), 2.0 ** 46, ctx.chain(5)), seed=900)
    srv = shm.BootstrapServer(boot, gk, rk, max_group=2)
    with pytest.raises(shm.capi.ShmInvalidArgument):
        srv.submit([bad])
    reqs = _requests(env, [1], 3)
    outs, _ = srv.wait(srv.submit(reqs[0][1]))
    _check_alone(env, reqs, [outs], tol_check=False)
    srv.close()


def test_server_two_workers_free_memory_admission(env):
    ctx, sk, boot, gk, rk = env
    sizes = [2, 1, 3, 1, 2, 2]
    reqs = _requests(env, sizes, 4)
    srv = shm.BootstrapServer(boot, gk, rk, mem_budget=0, max_group=2, workers=2)
    ids = [srv.submit(cts) for _, cts in reqs]
    results = [srv.wait(i)[0] for i in ids]
    st = srv.stats()
    print("two workers", st, flush=True)
    assert st["workers"] == 2 and st["completed"] == len(sizes) and st["in_flight_bytes"] == 0
    assert st["free_bytes"] > 0 and st["admit_cap_bytes"] > 0
    _check_alone(env, reqs, results)
    srv.close()


@pytest.fixture(scope="module")
def packed_env(env):
    ctx, sk, boot, gk, rk = env
    n = ctx.N // 2
    steps = list(boot.rotation_steps())
    p = 16
    while p < n:
        steps += [p, -p]
        p *= 2
    gk2 = sk.create_galois_keys_from_steps(sorted(set(steps)), seed=24)
    return gk2


def test_server_packed_requests(env, packed_env):
    ctx, sk, boot, gk, rk = env
    gk2 = packed_env
    enc, dec = shm.Encoder(ctx), shm.Decryptor(ctx, sk)
    n = ctx.N // 2
    rng = np.random.default_rng(31)
    lens = [100, 16, 37, 128, 64]      # token counts; blocks 128, 16, 64, 128, 64
    reqs = []
    for r, L in enumerate(lens):
        zs = []
        cts = []
        for col in range(2):
            z = np.zeros(n, complex)
            z[:L] = rng.uniform(-1, 1, L) + 1j * rng.uniform(-1, 1, L)
            zs.append(z)
            cts.append(sk.encrypt_symmetric(enc.encode(z, 2.0 ** 46, ctx.chain(4)), seed=700 + 10 * r + col))
        reqs.append((L, zs, cts))
    srv = shm.BootstrapServer(boot, gk2, rk, mem_budget=64 << 30, max_group=8, workers=1)
    # an ordinary request keeps the one worker busy while the packed ones queue up
    busy = _requests(env, [4], 7)
    bid = srv.submit(busy[0][1])
    ids = [srv.submit_packed(cts, L) for L, _, cts in reqs]
    res = [srv.wait(i) for i in ids]
    srv.wait(bid)
    st = srv.stats()
    offs = [r[2] for r in res]
    print("packed", st, "offsets", offs, flush=True)
    assert st["packed_requests"] == len(lens) and st["failed"] == 0
    blocks = [1 << (L - 1).bit_length() for L in lens]
    for o, b in zip(offs, blocks):
        assert o % b == 0 and o + b <= n
    packed_cts = st["ciphertexts"] - 4
    assert packed_cts < 2 * len(lens)  # fewer bootstraps than requests x columns
    if packed_cts == 2:  # one pack: the blocks are disjoint
        spans = sorted((o, o + b) for o, b in zip(offs, blocks))
        assert all(spans[i][1] <= spans[i + 1][0] for i in range(len(spans) - 1))
    for (L, zs, cts), (outs, lat, off) in zip(reqs, res):
        blk = 1 << (L - 1).bit_length()
        # alone: the same request bootstrapped by itself (offset 0)
        alone = [boot.bootstrap_s2c_first([c], gk2, rk)[0] for c in cts]
        for col in range(2):
            got = enc.decode(dec.decrypt(outs[col]), complex_out=True)
            ref = enc.decode(dec.decrypt(alone[col]), complex_out=True)
            assert float(np.max(np.abs(got[off:off + L] - zs[col][:L]))) <= 2e-5
            assert float(np.max(np.abs(got[off:off + L] - ref[:L]))) <= 2e-5
            if L < blk:
                assert float(np.max(np.abs(got[off + L:off + blk]))) <= 2e-5
    srv.close()


def test_server_packed_needs_keys(env):
    ctx, sk, boot, gk, rk = env
    reqs = _requests(env, [1], 6)
    srv = shm.BootstrapServer(boot, gk, rk, max_group=2)
    with pytest.raises(shm.capi.ShmInvalidArgument):
        srv.submit_packed(reqs[0][1], 64)
    srv.close()
