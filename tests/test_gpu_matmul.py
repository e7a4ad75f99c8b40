"""GPU parity of the MOAI ct-pt matmul (Ct_pt_matrix_mul.cuh:18-185).

- small shape: every variant bit-exact against the oracle on all outputs,
  at row_W = 20 (one launch) and row_W = 300 (two row chunks);
- row_W = 768 (the BERT projection width, three row chunks): INT8-MFMA and
  64-bit paths bit-exact against the oracle;
- full quarter-width shape (N=2^16, L=24, 192 x 192): INT8-MFMA == 64-bit
  path on every limb of every output, and sampled output columns bit-exact
  against the oracle;
- masked form (Ct_pt_matrix_mul.cuh:846-985): bit-exact against the oracle
  composed from its encode (vector), multiply_plain, add and rescale, over two
  row blocks and two output blocks; with an all-ones mask and a power-of-two
  scale it equals the unmasked matmul limb for limb (the inverse special FFT of
  a constant vector is exact);
- closed-form known answer from the reference's own test
  (test_ct_pt_matrix_mul.cuh: X[i][j][k] = j+1, W = 1/128 -> slot of token j
  decodes to (row_W/128) * (j+1)).
"""
import numpy as np
import pytest

import shmoai_amd as shm

pytestmark = pytest.mark.gpu


def _rand_cts(ctx, ref, rng, n, Ql):
    xs, hs = [], []
    for _ in range(n):
        a = np.empty((2, Ql, ref.N), np.uint64)
        for k in range(Ql):
            a[:, k, :] = rng.integers(0, int(ref.primes[k]), size=(2, ref.N), dtype=np.uint64)
        a = a.ravel()
        ct = shm.Ciphertext(ctx, 2, ctx.chain(Ql)).load(a)
        ct.scale = 2.0 ** 40
        xs.append(a)
        hs.append(ct)
    return xs, hs


@pytest.mark.parametrize("row_w", [20, 300])
@pytest.mark.parametrize("variant", [shm.MATMUL_U64, shm.MATMUL_INT8, shm.MATMUL_PER_CT])
def test_matmul_small_bitexact(oracle_lib, variant, row_w):
    bits = [50] * 5 + [60]
    ctx = shm.Context(13, bits, 1)
    ref = oracle_lib.Ref(13, bits, 1)
    rng = np.random.default_rng(5)
    col_w, Ql = 12, 5
    xs, hs = _rand_cts(ctx, ref, rng, row_w, Ql)
    W = np.random.default_rng(6).normal(0, 0.02, (row_w, col_w))
    out = shm.ct_pt_matrix_mul_wo_pre(hs, W, row_w, col_w, row_w, ctx, variant=variant)
    assert shm.last_matmul_variant() == variant
    exp = ref.ctpt_matmul(xs, W, Ql, np.full(row_w, 2.0 ** 40))
    for i in range(col_w):
        assert out[i].chain_index == ctx.chain(Ql - 1)
        assert out[i].scale == 2.0 ** 40
        assert np.array_equal(out[i].numpy(), exp[i]), f"output {i}"


def test_matmul_row_chunks_768(oracle_lib):
    bits = [50] * 4 + [60]
    ctx = shm.Context(13, bits, 1)
    ref = oracle_lib.Ref(13, bits, 1)
    rng = np.random.default_rng(7)
    row_w, col_w, Ql = 768, 8, 4
    xs, hs = _rand_cts(ctx, ref, rng, row_w, Ql)
    W = np.random.default_rng(8).normal(0, 0.02, (row_w, col_w))
    W[5, 3], W[700, 0] = 0.45, -0.45  # near the 5-digit weight bound, first and last chunk
    exp = ref.ctpt_matmul(xs, W, Ql, np.full(row_w, 2.0 ** 40))
    for variant in (shm.MATMUL_INT8, shm.MATMUL_U64):
        out = shm.ct_pt_matrix_mul_wo_pre(hs, W, row_w, col_w, row_w, ctx, variant=variant)
        assert shm.last_matmul_variant() == variant
        for i in range(col_w):
            assert np.array_equal(out[i].numpy(), exp[i]), f"variant {variant} output {i}"


def test_matmul_quarter_width_int8_equals_u64_and_oracle(oracle_lib):
    bits = [50] * 24 + [60]
    ctx = shm.Context(16, bits, 1)
    ref = oracle_lib.Ref(16, bits, 1)
    rng = np.random.default_rng(5)
    row_w = col_w = 192
    Ql = 24
    xs, hs = _rand_cts(ctx, ref, rng, row_w, Ql)
    W = np.random.default_rng(6).normal(0, 0.02, (row_w, col_w))
    o8 = shm.ct_pt_matrix_mul_wo_pre(hs, W, row_w, col_w, row_w, ctx, variant=shm.MATMUL_INT8)
    assert shm.last_matmul_variant() == shm.MATMUL_INT8
    o64 = shm.ct_pt_matrix_mul_wo_pre(hs, W, row_w, col_w, row_w, ctx, variant=shm.MATMUL_U64)
    for i in range(col_w):
        assert np.array_equal(o8[i].numpy(), o64[i].numpy()), f"output {i}"
    oracle_lib.set_threads(16)
    for i in (0, 191):
        exp = ref.ctpt_matmul(xs, W, Ql, np.full(row_w, 2.0 ** 40), i, i + 1)[0]
        assert np.array_equal(o8[i].numpy(), exp), f"output {i}"


LAYER_BITS = [51] + [46] * 23 + [51] * 11 + [51] * 5   # test_single_layer.cuh:473-535


@pytest.mark.parametrize("logN,Ql,row_w,col_w,oracle_cols", [(13, 6, 40, 12, None), (16, 24, 192, 192, (0, 191))])
def test_matmul_layer_primes_int8_equals_u64_and_oracle(oracle_lib, logN, Ql, row_w, col_w, oracle_cols):
    """The reference layer's prime chain (a 51-bit prime, then 46-bit primes) at scale 2^46:
    the INT8-MFMA contraction takes 46-bit primes (7 limb digits, 6 weight digits) and
    equals the 64-bit path and the oracle."""
    ctx = shm.Context(logN, LAYER_BITS, 5)
    ref = oracle_lib.Ref(logN, LAYER_BITS, 5)
    rng = np.random.default_rng(15)
    xs, hs = _rand_cts(ctx, ref, rng, row_w, Ql)
    for h in hs:
        h.scale = 2.0 ** 46
    W = np.random.default_rng(16).normal(0, 0.02, (row_w, col_w))
    o8 = shm.ct_pt_matrix_mul_wo_pre(hs, W, row_w, col_w, row_w, ctx, variant=shm.MATMUL_INT8)
    assert shm.last_matmul_variant() == shm.MATMUL_INT8
    o64 = shm.ct_pt_matrix_mul_wo_pre(hs, W, row_w, col_w, row_w, ctx, variant=shm.MATMUL_U64)
    for i in range(col_w):
        assert np.array_equal(o8[i].numpy(), o64[i].numpy()), f"output {i}"
    oracle_lib.set_threads(16)
    cols = range(col_w) if oracle_cols is None else oracle_cols
    for i in cols:
        exp = ref.ctpt_matmul(xs, W, Ql, np.full(row_w, 2.0 ** 46), i, i + 1)[0]
        assert np.array_equal(o8[i].numpy(), exp), f"output {i}"


def test_matmul_known_answer():
    """X[i][j][k] = j+1 (token j), W = 1/128: every slot of token j decodes to (row_W/128)(j+1)."""
    bits = [60] + [40] * 4 + [60]
    ctx = shm.Context(13, bits, 1)
    n_slots = ctx.N // 2
    num_X, num_row, row_w, col_w = 32, 128, 64, 8
    assert num_X * num_row == n_slots
    # batch_input packing (Batch_encode_encrypt.cuh:12-43): slot[num_X*k + j] = X[j][k][i]
    slot = np.empty(n_slots)
    for k in range(num_row):
        slot[num_X * k:num_X * (k + 1)] = k + 1.0
    sk = shm.SecretKey(ctx, seed=3, hamming_weight=64)
    enc = shm.Encoder(ctx)
    cts = [sk.encrypt_symmetric(enc.encode(slot, 2.0 ** 40, 1), seed=100 + i) for i in range(row_w)]
    W = np.full((row_w, col_w), 1.0 / 128.0)
    dec = shm.Decryptor(ctx, sk)
    for variant in (shm.MATMUL_INT8, shm.MATMUL_U64):
        out = shm.ct_pt_matrix_mul_wo_pre(cts, W, row_w, col_w, row_w, ctx, variant=variant)
        for o in out[:2]:
            v = enc.decode(dec.decrypt(o))
            expect = np.repeat(np.arange(1, num_row + 1) * (row_w / 128.0), num_X)
            assert np.max(np.abs(v - expect)) < 1e-3


def test_matmul_masked_bitexact(oracle_lib):
    bits = [50] * 4 + [60]
    ctx = shm.Context(13, bits, 1)
    ref = oracle_lib.Ref(13, bits, 1)
    rng = np.random.default_rng(9)
    row_w, col_w, Ql = 10, 20, 4
    xs, hs = _rand_cts(ctx, ref, rng, row_w, Ql)
    W = np.random.default_rng(10).normal(0, 0.02, (row_w, col_w))
    mask = (rng.random(ctx.N // 2) < 0.6).astype(np.int32)
    mask[:7] = [1, 0, 2, 1, -1, 1, 0]  # only entries equal to 1 select (bias_vec[j] == 1)
    out = shm.ct_pt_matrix_mul_wo_pre_w_mask(hs, W, mask, row_w, col_w, row_w, ctx)
    exp = ref.ctpt_matmul_masked(xs, W, mask, Ql, np.full(row_w, 2.0 ** 40))
    for i in range(col_w):
        assert out[i].chain_index == ctx.chain(Ql - 1)
        assert out[i].scale == 2.0 ** 40
        assert np.array_equal(out[i].numpy(), exp[i]), f"output {i}"


def test_matmul_masked_all_ones_equals_unmasked(oracle_lib):
    bits = [50] * 6 + [60]
    ctx = shm.Context(13, bits, 1)
    ref = oracle_lib.Ref(13, bits, 1)
    rng = np.random.default_rng(11)
    row_w, col_w, Ql = 24, 18, 6
    xs, hs = _rand_cts(ctx, ref, rng, row_w, Ql)
    W = np.random.default_rng(12).normal(0, 0.02, (row_w, col_w))
    om = shm.ct_pt_matrix_mul_wo_pre_w_mask(hs, W, np.ones(ctx.N // 2, np.int32), row_w, col_w, row_w, ctx)
    ou = shm.ct_pt_matrix_mul_wo_pre(hs, W, row_w, col_w, row_w, ctx, variant=shm.MATMUL_U64)
    for i in range(col_w):
        assert np.array_equal(om[i].numpy(), ou[i].numpy()), f"output {i}"


def test_matmul_masked_rejects_short_mask():
    ctx = shm.Context(13, [50] * 3 + [60], 1)
    hs = [shm.Ciphertext(ctx, 2, 1) for _ in range(2)]
    with pytest.raises(shm.capi.ShmInvalidArgument):
        shm.ct_pt_matrix_mul_wo_pre_w_mask(hs, np.zeros((2, 2)), np.ones(10, np.int32), 2, 2, 2, ctx)


def test_matmul_overlap_equals_serial(oracle_lib):
    """The INT8 matmul with its contraction in limb windows overlapped with the rescale on
    the context's second stream gives the limbs of the serial form (N = 2^13, 6 limbs,
    40 x 12 weights); both equal the oracle."""
    bits = [50] * 6 + [60]
    ctx = shm.Context(13, bits, 1)
    ref = oracle_lib.Ref(13, bits, 1)
    rng = np.random.default_rng(17)
    row_w, col_w, Ql = 40, 12, 6
    xs, hs = _rand_cts(ctx, ref, rng, row_w, Ql)
    W = np.random.default_rng(18).normal(0, 0.02, (row_w, col_w))
    try:
        shm.set_tuning("matmul_overlap", 0)
        serial = [o.numpy() for o in shm.ct_pt_matrix_mul_wo_pre(hs, W, row_w, col_w, row_w, ctx)]
        shm.set_tuning("matmul_overlap", 1)
        over = [o.numpy() for o in shm.ct_pt_matrix_mul_wo_pre(hs, W, row_w, col_w, row_w, ctx)]
    finally:
        shm.set_tuning("matmul_overlap", 0)
    assert shm.last_matmul_variant() == shm.MATMUL_INT8
    exp = ref.ctpt_matmul(xs, W, Ql, np.full(row_w, 2.0 ** 40))
    for i in range(col_w):
        assert np.array_equal(serial[i], over[i]), f"output {i}"
        assert np.array_equal(over[i], exp[i]), f"output {i}"
