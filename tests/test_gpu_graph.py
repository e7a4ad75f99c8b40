"""Graph capture of evaluator calls on a created stream (the enqueue-only contract of
the C ABI): copy, multiply_plain, rescale, rotate (key switch with temporaries
allocated inside the graph) and add are recorded once and replayed three times;
each replay, with the work ciphertext cleared and other default-stream work in
between, gives exactly the limbs of the eager run."""
import numpy as np
import pytest

import shmoai_amd as shm

pytestmark = pytest.mark.gpu

BITS = [50] * 5 + [60]


def test_graph_replay_equals_eager():
    ctx = shm.Context(13, BITS, 1)
    sk = shm.SecretKey(ctx, seed=5, hamming_weight=64)
    gk = sk.create_galois_keys_from_steps([1], seed=6)
    enc, ev = shm.Encoder(ctx), shm.Evaluator(ctx)
    rng = np.random.default_rng(7)
    n = ctx.N // 2
    X = sk.encrypt_symmetric(enc.encode(rng.uniform(-1, 1, n), 2.0 ** 40, 1), seed=8)
    Y = sk.encrypt_symmetric(enc.encode(rng.uniform(-1, 1, n), 2.0 ** 40, 2), seed=9)
    pt = enc.encode(rng.uniform(-1, 1, n), 2.0 ** 40, 1)
    W = shm.Ciphertext(ctx, 2, 1)
    st = shm.Stream()
    L = shm.capi.load()

    def sequence():
        shm.check(L.shm_ciphertext_copy(X.h, W.h, st.h))
        ev.multiply_plain_inplace(W, pt, stream=st.h)
        ev.rescale_to_next_inplace(W, stream=st.h)
        W.scale = Y.scale  # host metadata: the reference's callers align scales this way
        ev.rotate_vector_inplace(W, 1, gk, stream=st.h)
        ev.add_inplace(W, Y, stream=st.h)

    sequence()  # eager run: builds the Galois table and moves the buffers onto the stream
    shm.synchronize(st.h)
    expect = W.numpy()
    assert W.chain_index == 2

    g = shm.Graph(st)
    g.begin()
    sequence()
    g.end()
    spare = Y.copy()
    for rep in range(3):
        W.load(np.zeros_like(expect))
        g.launch()
        ev.rotate_vector_inplace(spare, 1, gk)  # unrelated default-stream work
        shm.synchronize(st.h)
        shm.synchronize(None)
        assert np.array_equal(W.numpy(), expect), f"replay {rep}"
    g.close()
    st.close()
