// Public-key encryption and the encrypted-input packing of the MOAI matmul.
//
// gen_public_key      secretkey.cu:524-535 gen_publickey (encrypt_zero_symmetric at QP)
// encrypt_asymmetric  secretkey.cu:145-217 encrypt_asymmetric (CKKS), with
//                     encrypt_zero_asymmetric_internal(_internal) (secretkey.cu:15-143)
// batch_input         Batch_encode_encrypt.cuh:12-43 (interleaved column packing, then
//                     encode and encrypt every column)
//
// Randomness is injected: a key or ciphertext draws its polynomials from
// splitmix64(seed, k) streams, k = 0 for the uniform (public key) or ternary
// (encryption) polynomial and k = 1 for the error.
#include <algorithm>
#include <vector>

#include "engine.h"

namespace shm {

namespace {

size_t qp_words(const Context &c) { return c.nQP() * c.N; }

// NTT over all QP limbs of one polynomial (data primes, then special primes)
void ntt_qp(const Context &c, uint64_t *p, hipStream_t s) {
    ntt_forward(c, p, 1, c.nQ, 0, s);
    ntt_forward(c, p + c.nQ * c.N, 1, c.nP, c.nQ, s);
}

}  // namespace

void gen_public_key(Context &c, const SecretKey &sk, uint64_t seed, PublicKey &out, hipStream_t s) {
    const size_t S = qp_words(c);
    out.ctx = &c;
    out.pk.alloc(2 * S * sizeof(uint64_t));
    uint64_t *p0 = out.pk.as<uint64_t>(), *p1 = p0 + S;
    Scratch e(S * sizeof(uint64_t), s);
    launch_sample_uniform(c, host_splitmix64(seed, 0), p1, c.nQ, true, s);
    launch_sample_gaussian_rns(c, host_splitmix64(seed, 1), e.as<uint64_t>(), c.nQ, true, s);
    ntt_qp(c, e.as<uint64_t>(), s);
    launch_neg_as_plus_e(c, sk.s.as<uint64_t>(), p1, e.as<uint64_t>(), p0, s);
}

// The zero encryption is formed at the key level (chain index 0): u uniform ternary,
// c'_p = u pk_p + e with one error polynomial for both components (the reference
// samples both from the same seed bytes, secretkey.cu:48-68), then each component is
// brought to Q by the ModDown of rns_bconv.cu:712-761 and the plaintext is added to c_0.
// Only plaintexts at chain index 1 are accepted: for a lower level the reference runs
// the level-1 ModDown tool on a ciphertext of the previous data level
// (secretkey.cu:111-139), which is not a defined operation.
void encrypt_asymmetric(Context &c, const PublicKey &pk, const uint64_t *pt, size_t pt_chain, double pt_scale,
                        uint64_t seed, Ciphertext &out, hipStream_t s) {
    if (pk.ctx != &c) throw InvalidArgument("public key belongs to another context");
    if (pt_chain != 1) throw InvalidArgument("encrypt_asymmetric: the plaintext must be at chain index 1");
    const size_t S = qp_words(c), nQ = c.nQ, N = c.N;
    Scratch u(S * sizeof(uint64_t), s), e(S * sizeof(uint64_t), s), cx(2 * S * sizeof(uint64_t), s);
    launch_sample_ternary_uniform(c, host_splitmix64(seed, 0), u.as<uint64_t>(), nQ, true, s);
    launch_sample_gaussian_rns(c, host_splitmix64(seed, 1), e.as<uint64_t>(), nQ, true, s);
    ntt_qp(c, u.as<uint64_t>(), s);
    ntt_qp(c, e.as<uint64_t>(), s);
    launch_pk_mac(c, u.as<uint64_t>(), pk.pk.as<uint64_t>(), e.as<uint64_t>(), cx.as<uint64_t>(), 2, s);
    out.ctx = &c;
    out.resize(2, 1, s, false);
    out.ntt = true;
    ks_moddown(c, cx.as<uint64_t>(), 2, nQ, out.data(), nQ * N, 0, s, false);
    launch_add(c, out.poly(0), pt, nQ, 1, s);
    out.scale = pt_scale;
}

void batch_input(Context &c, const PublicKey &pk, const double *X, size_t num_x, size_t num_row, size_t num_col,
                 double scale, uint64_t seed, const std::vector<Ciphertext *> &out, hipStream_t s) {
    const size_t N = c.N, n = N / 2, nQ = c.nQ;
    if (out.size() != num_col) throw InvalidArgument("batch_input: one output ciphertext per column");
    if (num_x == 0 || num_row == 0) throw InvalidArgument("batch_input: empty input");
    if (num_x * num_row > n) throw InvalidArgument("batch_input: num_X * num_row exceeds the slot count");
    if (!(scale > 0)) throw InvalidArgument("scale out of bounds");
    if (num_col == 0) return;
    const size_t words = num_x * num_row * num_col;
    Scratch dX(words * sizeof(double), s);
    SHM_HIP(hipMemcpyAsync(dX.as<void>(), X, words * sizeof(double), hipMemcpyHostToDevice, s));
    // columns are encoded in blocks: B slot vectors (B x N/2 complex) and B plaintexts
    const size_t B = std::min<size_t>(num_col, 32);
    std::vector<double> fixh(B, scale / (double)n);
    Scratch fix(B * sizeof(double), s), fft(B * n * sizeof(double2), s), pts(B * nQ * N * sizeof(uint64_t), s);
    SHM_HIP(hipMemcpyAsync(fix.as<void>(), fixh.data(), B * sizeof(double), hipMemcpyHostToDevice, s));
    for (size_t c0 = 0; c0 < num_col; c0 += B) {
        const size_t nb = std::min(B, num_col - c0);
        encode_columns_batch(c, dX.as<double>(), num_x, num_row, num_col, c0, nb, nQ, fix.as<double>(),
                             pts.as<uint64_t>(), fft.as<double2>(), s);
        for (size_t b = 0; b < nb; b++)
            encrypt_asymmetric(c, pk, pts.as<uint64_t>() + b * nQ * N, 1, scale, host_splitmix64(seed, c0 + b),
                               *out[c0 + b], s);
    }
    // the host copy of X may be released once the upload has finished
    SHM_HIP(hipStreamSynchronize(s));
}

}  // namespace shm
