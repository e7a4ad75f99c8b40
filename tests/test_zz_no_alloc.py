"""The allocation-free rescale (shm_rescale_to_next_inplace_noalloc) against the
allocating entry point, limb for limb, through the C ABI: two calls in a row, the
second with a smaller ciphertext (the static scratch array keeps the first call's
words), plus the oracle's floor rescale and the refusal of a ciphertext whose
scratch would exceed the array."""
import numpy as np
import pytest

import shmoai_amd as shm

pytestmark = pytest.mark.gpu


def _rand(ctx, rng, npolys, Ql):
    out = np.empty((npolys, Ql, ctx.N), np.uint64)
    for k in range(Ql):
        out[:, k, :] = rng.integers(0, int(ctx.coeff_modulus[k]), size=(npolys, ctx.N), dtype=np.uint64)
    return out.ravel()


def test_noalloc_rescale_equals_allocating_rescale(oracle_lib):
    ctx = shm.Context(16, [50] * 24 + [60], 1)
    ev = shm.Evaluator(ctx)
    rng = np.random.default_rng(71)
    # first call: 2 polynomials x 24 limbs; second call: 2 polynomials x 9 limbs
    for Ql in (24, 9):
        x = _rand(ctx, rng, 2, Ql)
        a = shm.Ciphertext(ctx, 2, ctx.chain(Ql)).load(x)
        b = shm.Ciphertext(ctx, 2, ctx.chain(Ql)).load(x)
        a.scale = b.scale = 2.0 ** 80
        ev.rescale_to_next_inplace(a)
        ev.rescale_to_next_inplace_noalloc(b)
        assert b.chain_index == a.chain_index == ctx.chain(Ql - 1)
        assert b.scale == a.scale
        got = b.numpy()
        assert np.array_equal(got, a.numpy()), f"Ql={Ql}"
        if Ql == 9:
            ref = oracle_lib.Ref(16, [50] * 24 + [60], 1)
            assert np.array_equal(got, ref.rescale(x, Ql))


def test_noalloc_rescale_small_ring_after_large():
    """A ring of another degree after the first test's calls: N = 2^13, three
    polynomials (a tensor product before relinearisation)."""
    ctx = shm.Context(13, [50] * 8 + [60], 1)
    ev = shm.Evaluator(ctx)
    x = _rand(ctx, np.random.default_rng(72), 3, 8)
    a = shm.Ciphertext(ctx, 3, 1).load(x)
    b = shm.Ciphertext(ctx, 3, 1).load(x)
    ev.rescale_to_next_inplace(a)
    ev.rescale_to_next_inplace_noalloc(b)
    assert np.array_equal(b.numpy(), a.numpy())


def test_noalloc_rescale_refuses_oversized():
    ctx = shm.Context(16, [40] * 44 + [60], 1)
    ct = shm.Ciphertext(ctx, 3, 1)  # 3 x 44 x 2^16 words > 8 Mi words
    with pytest.raises(shm.capi.ShmInvalidArgument):
        shm.Evaluator(ctx).rescale_to_next_inplace_noalloc(ct)
    assert ct.chain_index == 1
