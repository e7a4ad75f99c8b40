// Operations of the moai::Evaluator facade that compose the evaluator primitives
// (ckks_evaluator_parallel.cuh:444-561, ckks_evaluator_parallel.cu:466-590).
//
// Every constant goes through the full-slot encoder exactly as the facade does:
// encode(vector(slot_count, value), ct.scale) at chain index 1, mod-switched to the
// ciphertext's level. A limb of an encoded plaintext depends only on its own prime,
// so encoding straight at the ciphertext's level gives the same limbs.
#include <cmath>

#include "engine.h"

namespace shm {

namespace {

void need_ntt(const Ciphertext &ct) {
    if (!ct.ntt) throw InvalidArgument("CKKS encrypted must be in NTT form");
}

// rescale_to_next_inplace (evaluate.cu:1563-1584, floor division)
void rescale_floor(Context &c, Ciphertext &ct, hipStream_t s) { rescale_inplace(c, ct, s, false); }

// multiply_const (ckks_evaluator_parallel.cuh:444-460) into a fresh ciphertext
void multiply_const_copy(Context &c, const Ciphertext &in, double value, Ciphertext &out, hipStream_t s) {
    ct_clone(out, in, s);
    facade_multiply_const(c, out, value, 0.0, s);
}

// the tensor product of multiply_inplace (evaluate.cu:350-402): dst = a * b (size 3)
void tensor_into(Context &c, const Ciphertext &a, const Ciphertext &b, Ciphertext &dst, hipStream_t s) {
    if (a.size != 2 || b.size != 2) throw InvalidArgument("multiply supports size-2 ciphertexts");
    if (a.chain != b.chain) throw InvalidArgument("encrypted1 and encrypted2 parameter mismatch");
    const size_t Ql = a.Ql(), S3 = 3 * Ql * c.N;
    Scratch t(S3 * sizeof(uint64_t), s);
    if (&a == &b)
        launch_square(c, a.data(), t.as<uint64_t>(), Ql, s);
    else
        launch_tensor(c, a.data(), b.data(), t.as<uint64_t>(), Ql, s);
    const double scale = a.scale * b.scale;
    dst.ctx = &c;
    dst.resize(3, a.chain, s, false);
    launch_copy_words(dst.data(), t.as<uint64_t>(), S3, s);
    dst.ntt = true;
    dst.scale = scale;
}

}  // namespace

// multiply_const_inplace (ckks_evaluator_parallel.cuh:451-460): ct *= encode(value) at ct.scale
void facade_multiply_const(Context &c, Ciphertext &ct, double re, double im, hipStream_t s) {
    need_ntt(ct);
    const size_t Ql = ct.Ql();
    Scratch pt(Ql * c.N * sizeof(uint64_t), s);
    encode_const(c, re, im, Ql, ct.scale, pt.as<uint64_t>(), s);
    launch_mul_plain(c, ct.data(), pt.as<uint64_t>(), Ql, ct.size, s);
    ct.scale = ct.scale * ct.scale;
}

// add_const_inplace (ckks_evaluator_parallel.cuh:469-478): c_0 += encode(value) at ct.scale
void facade_add_const(Context &c, Ciphertext &ct, double re, double im, hipStream_t s) {
    need_ntt(ct);
    const size_t Ql = ct.Ql();
    Scratch pt(Ql * c.N * sizeof(uint64_t), s);
    encode_const(c, re, im, Ql, ct.scale, pt.as<uint64_t>(), s);
    launch_add(c, ct.poly(0), pt.as<uint64_t>(), Ql, 1, s);
}

// multiply_vector_inplace_reduced_error (ckks_evaluator_parallel.cuh:538-560): the
// vector, padded with zeros to the slot count, encoded at ct.scale, then multiply_plain
void facade_multiply_vector(Context &c, Ciphertext &ct, const double *re, const double *im, size_t n,
                            hipStream_t s) {
    need_ntt(ct);
    const size_t Ql = ct.Ql();
    Scratch pt(Ql * c.N * sizeof(uint64_t), s);
    encode(c, re, im, n, Ql, ct.scale, pt.as<uint64_t>(), s);
    launch_mul_plain(c, ct.data(), pt.as<uint64_t>(), Ql, ct.size, s);
    ct.scale = ct.scale * ct.scale;
}

// add/sub/multiply_inplace_reduced_error (ckks_evaluator_parallel.cu:466-590). Operands
// at one level: ct1 takes ct2's scale, then the plain operation. Otherwise the operand
// with more limbs is multiplied by the constant s_low * q_last / s_high^2, labelled
// s_low * q_last, rescaled (to s_low) and mod-switched to the other's level; the result
// carries the lower operand's scale. Multiply relinearises in every case.
void facade_reduced_error(Context &c, Ciphertext &ct1, const Ciphertext &ct2, int op, const KSwitchKey *rk,
                          hipStream_t s) {
    if (op < 0 || op > 2) throw InvalidArgument("reduced_error: unknown operation");
    if (op == 2 && !rk) throw InvalidArgument("reduced_error multiply needs relinearisation keys");
    need_ntt(ct1);
    need_ntt(ct2);
    if (ct1.size != 2 || ct2.size != 2) throw InvalidArgument("reduced_error operations take size-2 ciphertexts");
    const size_t Q1 = ct1.Ql(), Q2 = ct2.Ql();
    auto plain_op = [&](Ciphertext &dst, const Ciphertext &a, const Ciphertext &b) {
        // dst = a (op) b; dst may be a
        if (op == 2) {
            tensor_into(c, a, b, dst, s);
            return;
        }
        if (&dst != &a) ct_clone(dst, a, s);
        if (op == 0)
            launch_add(c, dst.data(), b.data(), dst.Ql(), 2, s);
        else
            launch_sub(c, dst.data(), b.data(), dst.Ql(), 2, s);
    };
    if (Q1 == Q2) {
        ct1.scale = ct2.scale;
        plain_op(ct1, ct1, ct2);
    } else if (Q1 < Q2) {
        const double qlast = (double)c.q[Q2 - 1];
        const double adjust = ct1.scale * qlast / (ct2.scale * ct2.scale);
        Ciphertext adj;
        multiply_const_copy(c, ct2, adjust, adj, s);
        adj.scale = ct1.scale * qlast;
        rescale_floor(c, adj, s);
        mod_switch_drop(c, adj, ct1.chain, s);
        ct1.scale = adj.scale;
        plain_op(ct1, ct1, adj);
    } else {
        const double qlast = (double)c.q[Q1 - 1];
        const double adjust = ct2.scale * qlast / (ct1.scale * ct1.scale);
        Ciphertext adj;
        multiply_const_copy(c, ct1, adjust, adj, s);
        adj.scale = ct2.scale * qlast;
        rescale_floor(c, adj, s);
        mod_switch_drop(c, adj, ct2.chain, s);
        adj.scale = ct2.scale;
        plain_op(ct1, adj, ct2);
    }
    if (op == 2) {
        keyswitch(c, ct1.poly(2), rk->k.as<uint64_t>(), ct1.Ql(), ct1.poly(0), ct1.poly(1), true, s);
        ct1.size = 2;
    }
}

}  // namespace shm
