"""The C++ facade (include/shmoai.hpp), compiled with g++ against libshmoai.so
(tests/cpp/facade_c1, built by __graft_entry__.build()): the reference's facade call
sequence for the C1 pipeline and one ct_pt_matrix_mul_wo_pre give the oracle's limbs
bit for bit; errors arrive as std::invalid_argument; public-key encryption through
moai::Encryptor decrypts through moai::Decryptor."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "facade_c1")
C1_BITS = [50] * 8 + [60]


def test_cpp_facade_c1_pipeline_and_matmul(oracle_lib, tmp_path):
    assert os.path.exists(BIN), "tests/cpp/facade_c1 is not built (run __graft_entry__.build())"
    ref = oracle_lib.Ref(13, C1_BITS, 1)
    N, Ql = ref.N, 8
    rng = np.random.default_rng(3)

    def limbs(npolys):
        a = np.empty((npolys, Ql, N), np.uint64)
        for k in range(Ql):
            a[:, k, :] = rng.integers(0, int(ref.primes[k]), size=(npolys, N), dtype=np.uint64)
        return a.ravel()

    x = limbs(2)
    X = [limbs(2) for _ in range(4)]
    paths = [str(tmp_path / n) for n in ("ct_in", "ct_out", "mm_in", "mm_out")]
    x.tofile(paths[0])
    np.concatenate(X).tofile(paths[2])
    r = subprocess.run([BIN, *paths], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "FACADE_OK" in r.stdout, r.stdout + r.stderr
    # the C1 pipeline through the oracle, with the facade's full-slot encode of 0.37
    s_ref = ref.sk_ntt(ref.sample_ternary(1, 64))
    elt = ref.elt(1)
    key = ref.gen_galois_key(s_ref, elt, ref.splitmix64(9, elt))
    exp = ref.ntt(x, np.tile(ref.pidx(Ql), 2))
    exp = ref.rotate(exp, Ql, elt, key)
    exp = ref.multiply_plain(exp, ref.encode(np.full(N // 2, 0.37), Ql, 2.0 ** 50), Ql)
    exp = ref.rescale(exp, Ql)
    exp = ref.intt(exp, np.tile(ref.pidx(Ql - 1), 2))
    assert np.array_equal(np.fromfile(paths[1], np.uint64), exp)
    W = np.array([[0.01 * (j + 1) - 0.005 * i for i in range(3)] for j in range(4)])
    outs = ref.ctpt_matmul(X, W, Ql, np.full(4, 2.0 ** 40))
    assert np.array_equal(np.fromfile(paths[3], np.uint64), np.concatenate(outs))
