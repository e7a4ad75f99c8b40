"""Bootstrapping server with continuous batching (no counterpart in the reference).

- requests of uneven sizes (and one at another scale), submitted from two threads
  while others run, each get exactly the limbs of bootstrapping their ciphertexts
  alone, and decrypt within 2e-5;
- admission by memory: with a budget below one request, requests enter one at a
  time (so groups never mix requests) and the results are unchanged;
- a request with the wrong level is refused at submit and the server keeps serving.
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
