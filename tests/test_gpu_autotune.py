"""The inner-product autotuner (no counterpart in the reference): the first call of each
shape times the candidate member / output slot counts and caches the fastest. The
results must not depend on it: a grouped bootstrap with the autotuner on (cold cache)
equals the same bootstrap with one slot forced, limb for limb, and the report lists the
tuned shapes with their choice."""
import numpy as np
import pytest

import shmoai_amd as shm

pytestmark = pytest.mark.gpu

SMALL_BITS = [51] + [46] * 3 + [51] * 14 + [51] * 3


def test_autotuned_bootstrap_equals_forced():
    ctx = shm.Context(13, SMALL_BITS, 3)
    sk = shm.SecretKey(ctx, seed=31, hamming_weight=64)
    boot = shm.Bootstrapper(ctx)
    gk = sk.create_galois_keys_from_steps(boot.rotation_steps(), seed=32)
    rk = sk.gen_relinkey(seed=33)
    enc = shm.Encoder(ctx)
    rng = np.random.default_rng(3)
    n = ctx.N // 2
    cts = [sk.encrypt_symmetric(enc.encode(rng.uniform(-1, 1, n), 2.0 ** 46, ctx.chain(4)), seed=40 + i)
           for i in range(3)]
    try:
        shm.set_tuning("ks_batch_slots_log2", 0)
        shm.set_tuning("ks_multi_slots_log2", 0)
        forced = boot.bootstrap_s2c_first(cts, gk, rk)
        shm.set_tuning("ks_batch_slots_log2", -1)
        shm.set_tuning("ks_multi_slots_log2", -1)
        shm.autotune_clear()
        shm.autotune_enable(True)
        tuned = boot.bootstrap_s2c_first(cts, gk, rk)
        again = boot.bootstrap_s2c_first(cts, gk, rk)   # cached choices
        rep = shm.autotune_report()
        print(rep, flush=True)
        assert "batch N=8192" in rep and "-> slots" in rep
        for a, b, c in zip(forced, tuned, again):
            assert np.array_equal(a.numpy(), b.numpy())
            assert np.array_equal(a.numpy(), c.numpy())
    finally:
        shm.set_tuning("ks_batch_slots_log2", -1)
        shm.set_tuning("ks_multi_slots_log2", -1)
