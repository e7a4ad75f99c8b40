"""GPU checks of the Chebyshev polynomial evaluation used by the non-linear layers
(Polynomial::homomorphic_poly_evaluation, bootstrapping/common/Polynomial.cu:253-476;
gelu_cheb, non_linear_func/gelu_other.cuh:257-330).

- a degree-24 Chebyshev interpolant of GELU on [-19.5, 12.5] (the interval of
  gelu_cheb) and a degree-63 series with decaying random coefficients (giant step
  T_32) decrypt within 2e-5 of the FP64 evaluation of the same series;
- the output keeps the input scale and lands on the level the depth predicts;
- a batch of three equals three single calls limb for limb (members do not
  interact);
- degrees 9, 16, 32 and 48, where the split meets a power of two, evaluate at the
  predicted depth.
"""
import math

import numpy as np
import pytest
from numpy.polynomial import chebyshev as npc

import shmoai_amd as shm

pytestmark = pytest.mark.gpu

BITS = [60] + [40] * 10 + [60]
SCALE = 2.0 ** 40


def _gelu(x):
    return 0.5 * x * (1.0 + np.vectorize(math.erf)(x / math.sqrt(2.0)))


@pytest.fixture(scope="module")
def env():
    ctx = shm.Context(13, BITS, 1)
    sk = shm.SecretKey(ctx, seed=11, hamming_weight=64)
    rk = sk.gen_relinkey(seed=12)
    return ctx, sk, rk, shm.Encoder(ctx), shm.Decryptor(ctx, sk), shm.Evaluator(ctx)


def _encrypt(env, xs, seed):
    ctx, sk, _, enc, _, _ = env
    return sk.encrypt_symmetric(enc.encode(xs, SCALE, 1), seed=seed)


# gelu_cheb's own coefficients d_0..d_24 (non_linear_func/gelu_other.cuh:286-291), for
# sum_k d_k T_k(u), u = (x + 3.5) / 16 on [-19.5, 12.5]
GELU_CHEB_D = [
    3.4550752799e+00, 5.7852067632e+00, 3.1727714170e+00, 7.0234932500e-01, -4.6200569546e-01,
    -3.7817535966e-01, 8.4136420694e-02, 2.2427456987e-01, 2.9451916291e-02, -1.2928072669e-01,
    -6.6659430283e-02, 6.4349502364e-02, 7.1733295078e-02, -1.9882581810e-02, -6.1145844240e-02,
    -8.4827295421e-03, 4.3597254356e-02, 2.3373027950e-02, -2.4790187963e-02, -2.7940100474e-02,
    8.8404026343e-03, 2.3776567058e-02, 8.8171887083e-03, -4.6290425955e-02, -9.7997267900e-03]


@pytest.mark.parametrize("case", ["gelu24", "gelu_cheb_reference", "random63"])
def test_chebyshev_matches_fp64(env, case):
    ctx, sk, rk, enc, dec, ev = env
    n = ctx.N // 2
    rng = np.random.default_rng(21)
    if case == "gelu_cheb_reference":
        lo, hi = -19.5, 12.5
        coeffs = np.array(GELU_CHEB_D)
        depth = 6
    elif case == "gelu24":
        lo, hi = -19.5, 12.5
        coeffs = npc.chebinterpolate(lambda u: _gelu((u * (hi - lo) + lo + hi) / 2.0), 24)
        depth = 6
    else:
        lo, hi = -2.0, 3.0
        coeffs = rng.normal(0, 1, 64) / (1.0 + np.arange(64)) ** 1.5
        depth = 8
    x = rng.uniform(lo, hi, n)
    ct = _encrypt(env, x, 31)
    chain0 = ct.chain_index
    ev.evaluate_chebyshev_inplace(ct, coeffs, lo, hi, rk)
    assert ct.scale == SCALE
    assert ct.chain_index == chain0 + depth
    got = enc.decode(dec.decrypt(ct))
    want = npc.chebval((2.0 * x - lo - hi) / (hi - lo), coeffs)
    err = float(np.max(np.abs(got - want)))
    assert err < 2e-5, err
    if case == "gelu_cheb_reference":
        # the fit itself: gelu_cheb's stated maximum error 0.037 on the interval
        assert float(np.max(np.abs(want - _gelu(x)))) < 0.04


@pytest.mark.parametrize("deg,depth", [(9, 5), (16, 6), (32, 7), (48, 7)])
def test_chebyshev_split_degrees(env, deg, depth):
    """Degrees where a split meets a power of two (deg = 2m: T_m T_m = (T_2m + T_0) / 2)
    or leaves a quotient of degree at most 8; depth = 1 (affine map) + the recursion's."""
    ctx, sk, rk, enc, dec, ev = env
    rng = np.random.default_rng(100 + deg)
    lo, hi = -1.5, 2.5
    coeffs = rng.normal(0, 1, deg + 1) / (1.0 + np.arange(deg + 1)) ** 1.5
    x = rng.uniform(lo, hi, ctx.N // 2)
    ct = _encrypt(env, x, 60 + deg)
    chain0 = ct.chain_index
    ev.evaluate_chebyshev_inplace(ct, coeffs, lo, hi, rk)
    assert ct.scale == SCALE
    assert ct.chain_index == chain0 + depth
    want = npc.chebval((2.0 * x - lo - hi) / (hi - lo), coeffs)
    err = float(np.max(np.abs(enc.decode(dec.decrypt(ct)) - want)))
    assert err < 2e-5, err


def test_chebyshev_batch_equals_single(env):
    ctx, sk, rk, enc, dec, ev = env
    n = ctx.N // 2
    rng = np.random.default_rng(22)
    lo, hi = -4.0, 4.0
    coeffs = rng.normal(0, 1, 20) / (1.0 + np.arange(20)) ** 2
    xs = [rng.uniform(lo, hi, n) for _ in range(3)]
    batch = [_encrypt(env, x, 40 + i) for i, x in enumerate(xs)]
    single = [_encrypt(env, x, 40 + i) for i, x in enumerate(xs)]
    ev.evaluate_chebyshev_inplace(batch, coeffs, lo, hi, rk)
    for c in single:
        ev.evaluate_chebyshev_inplace(c, coeffs, lo, hi, rk)
    for i in range(3):
        assert np.array_equal(batch[i].numpy(), single[i].numpy()), f"member {i}"
        got = enc.decode(dec.decrypt(batch[i]))
        want = npc.chebval((2.0 * xs[i] - lo - hi) / (hi - lo), coeffs)
        assert float(np.max(np.abs(got - want))) < 2e-5


def test_chebyshev_rejects_bad_arguments(env):
    ctx, sk, rk, enc, dec, ev = env
    ct = _encrypt(env, np.zeros(ctx.N // 2), 50)
    with pytest.raises(shm.capi.ShmInvalidArgument):
        ev.evaluate_chebyshev_inplace(ct, [1.0], -1.0, 1.0, rk)        # constant only
    with pytest.raises(shm.capi.ShmInvalidArgument):
        ev.evaluate_chebyshev_inplace(ct, [0.0, 1.0], 1.0, 1.0, rk)    # empty interval
    with pytest.raises(shm.capi.ShmInvalidArgument):
        ev.evaluate_chebyshev_inplace(ct, np.ones(65), -1.0, 1.0, rk)  # degree 64
