/*
 * shmoai.hpp — C++ facade of libshmoai.so with the reference's class and method names.
 *
 * The reference's BERT composition layers call PhantomFHE types and the moai:: facade
 * (src/include/source/ckks_evaluator_parallel.cuh:13-680), the MOAI matmul
 * (matrix_mul/Ct_pt_matrix_mul.cuh:18-185), the input packing
 * (matrix_mul/Batch_encode_encrypt.cuh:12-57) and BatchBootstrapper
 * (bootstrapping/BatchBootstrap.cuh:470-820). This header gives the same names and
 * argument meaning over the C ABI (shmoai.h), minus the CUDA stream arguments (every
 * call runs on the library's default stream), so a call site such as
 *
 *     evaluator.multiply_plain_inplace(ct, pt);
 *     evaluator.rescale_to_next_inplace(ct);
 *     evaluator.rotate_vector_inplace(ct, 1, galois_keys);
 *
 * compiles unchanged against it. Differences a caller sees:
 *   - randomness is injected: keys and the Encryptor take a seed (the reference draws
 *     unseeded Salsa20 streams and an unseeded Hamming-weight shuffle);
 *   - ct.scale() returns an assignable proxy (the reference returns double&);
 *   - errors are C++ exceptions of the reference's classes, rethrown from the C status:
 *     std::invalid_argument, std::logic_error, std::bad_alloc, std::runtime_error.
 * Header-only; link with -lshmoai.
 */
#ifndef SHMOAI_HPP
#define SHMOAI_HPP

#include <complex>
#include <cstdint>
#include <new>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "shmoai.h"

namespace moai {

inline void shm_check(int rc) {
    if (rc == SHM_OK) return;
    const std::string msg = shm_last_error();
    switch (rc) {
        case SHM_ERR_INVALID_ARGUMENT:
            throw std::invalid_argument(msg);
        case SHM_ERR_LOGIC:
            throw std::logic_error(msg);
        case SHM_ERR_OUT_OF_MEMORY:
            throw std::bad_alloc();
        default:
            throw std::runtime_error(msg);
    }
}

// EncryptionParameters + CoeffModulus::Create + PhantomContext(parms) (context.cu:125-238):
// data primes then special primes, as passed to CoeffModulus::Create.
class PhantomContext {
  public:
    PhantomContext(uint32_t log_poly_degree, const std::vector<int> &coeff_bit_sizes, uint32_t special_modulus_size) {
        shm_check(shm_context_create(log_poly_degree, coeff_bit_sizes.data(), (uint32_t)coeff_bit_sizes.size(),
                                     special_modulus_size, &h_));
        shm_check(shm_context_get_info(h_, &info_));
    }
    ~PhantomContext() {
        if (h_) shm_context_destroy(h_);
    }
    PhantomContext(const PhantomContext &) = delete;
    PhantomContext &operator=(const PhantomContext &) = delete;
    shm_context *handle() const { return h_; }
    size_t poly_modulus_degree() const { return info_.poly_degree; }
    size_t size_q() const { return info_.size_q; }
    size_t size_p() const { return info_.size_p; }
    std::vector<uint64_t> coeff_modulus() const {
        std::vector<uint64_t> q(info_.size_q + info_.size_p);
        shm_check(shm_context_coeff_modulus(h_, q.data(), q.size()));
        return q;
    }
    uint32_t get_elt_from_step(int step) const {
        uint32_t e = 0;
        shm_check(shm_galois_elt_from_step(h_, step, &e));
        return e;
    }

  private:
    shm_context *h_ = nullptr;
    shm_context_info info_{};
};

namespace detail {
template <class Setter>
class ScaleRef {
  public:
    ScaleRef(double v, Setter set) : v_(v), set_(std::move(set)) {}
    operator double() const { return v_; }
    ScaleRef &operator=(double v) {
        set_(v);
        v_ = v;
        return *this;
    }

  private:
    double v_;
    Setter set_;
};
}  // namespace detail

// PhantomCiphertext (ciphertext.h:21-288): value semantics, copies are device copies.
class PhantomCiphertext {
  public:
    PhantomCiphertext() = default;
    PhantomCiphertext(const PhantomContext &ctx, uint32_t size = 2, uint32_t chain_index = 1) : ctx_(&ctx) {
        shm_check(shm_ciphertext_create(ctx.handle(), size, chain_index, &h_));
    }
    ~PhantomCiphertext() { release(); }
    PhantomCiphertext(const PhantomCiphertext &o) : ctx_(o.ctx_) {
        if (o.h_) {
            shm_check(shm_ciphertext_create(ctx_->handle(), o.size(), o.chain_index(), &h_));
            shm_check(shm_ciphertext_copy(o.h_, h_, nullptr));
        }
    }
    PhantomCiphertext &operator=(const PhantomCiphertext &o) {
        if (this != &o) {
            PhantomCiphertext t(o);
            swap(t);
        }
        return *this;
    }
    PhantomCiphertext(PhantomCiphertext &&o) noexcept { swap(o); }
    PhantomCiphertext &operator=(PhantomCiphertext &&o) noexcept {
        swap(o);
        return *this;
    }
    void swap(PhantomCiphertext &o) noexcept {
        std::swap(h_, o.h_);
        std::swap(ctx_, o.ctx_);
    }
    shm_ciphertext *handle() const { return h_; }
    const PhantomContext *context() const { return ctx_; }
    size_t size() const { return info().size; }
    size_t chain_index() const { return info().chain; }
    size_t coeff_modulus_size() const { return info().q; }
    size_t poly_modulus_degree() const { return ctx_->poly_modulus_degree(); }
    auto scale() {
        shm_ciphertext *h = h_;
        auto set = [h](double v) { shm_check(shm_ciphertext_set_scale(h, v)); };
        return detail::ScaleRef<decltype(set)>(info().scale, set);
    }
    double scale() const { return info().scale; }
    // host limbs [size][Ql][N]
    void load(const std::vector<uint64_t> &limbs) {
        shm_check(shm_ciphertext_import(h_, limbs.data(), limbs.size(), nullptr));
    }
    std::vector<uint64_t> limbs() const {
        std::vector<uint64_t> v(size() * coeff_modulus_size() * poly_modulus_degree());
        shm_check(shm_ciphertext_export(h_, v.data(), v.size(), nullptr));
        return v;
    }

  private:
    struct Info {
        uint32_t size, chain, q;
        double scale;
    };
    Info info() const {
        if (!h_) throw std::logic_error("empty ciphertext");
        Info i{};
        shm_check(shm_ciphertext_info(h_, &i.size, &i.chain, &i.q, &i.scale));
        return i;
    }
    void release() {
        if (h_) shm_ciphertext_destroy(h_);
        h_ = nullptr;
    }
    shm_ciphertext *h_ = nullptr;
    const PhantomContext *ctx_ = nullptr;
};

// PhantomPlaintext (plaintext.h): [Ql][N] limbs in NTT form.
class PhantomPlaintext {
  public:
    PhantomPlaintext() = default;
    PhantomPlaintext(const PhantomContext &ctx, uint32_t chain_index = 1) : ctx_(&ctx) {
        shm_check(shm_plaintext_create(ctx.handle(), chain_index, &h_));
    }
    ~PhantomPlaintext() {
        if (h_) shm_plaintext_destroy(h_);
    }
    PhantomPlaintext(const PhantomPlaintext &) = delete;
    PhantomPlaintext &operator=(const PhantomPlaintext &) = delete;
    PhantomPlaintext(PhantomPlaintext &&o) noexcept { swap(o); }
    PhantomPlaintext &operator=(PhantomPlaintext &&o) noexcept {
        swap(o);
        return *this;
    }
    void swap(PhantomPlaintext &o) noexcept {
        std::swap(h_, o.h_);
        std::swap(ctx_, o.ctx_);
    }
    shm_plaintext *handle() const { return h_; }
    size_t chain_index() const {
        uint32_t ch = 0, q = 0;
        double sc = 0;
        shm_check(shm_plaintext_info(h_, &ch, &q, &sc));
        return ch;
    }
    double scale() const {
        uint32_t ch = 0, q = 0;
        double sc = 0;
        shm_check(shm_plaintext_info(h_, &ch, &q, &sc));
        return sc;
    }
    // (re)create at a chain index when empty or at another level
    void ensure(const PhantomContext &ctx, uint32_t chain_index) {
        if (!h_ || chain_index_or0() != chain_index) {
            PhantomPlaintext t(ctx, chain_index);
            swap(t);
        }
    }

  private:
    size_t chain_index_or0() const { return h_ ? chain_index() : 0; }
    shm_plaintext *h_ = nullptr;
    const PhantomContext *ctx_ = nullptr;
};

class PhantomRelinKey {
  public:
    PhantomRelinKey() = default;
    explicit PhantomRelinKey(shm_kswitch_key *h) : h_(h) {}
    ~PhantomRelinKey() {
        if (h_) shm_kswitch_key_destroy(h_);
    }
    PhantomRelinKey(const PhantomRelinKey &) = delete;
    PhantomRelinKey &operator=(const PhantomRelinKey &) = delete;
    PhantomRelinKey(PhantomRelinKey &&o) noexcept : h_(o.h_) { o.h_ = nullptr; }
    PhantomRelinKey &operator=(PhantomRelinKey &&o) noexcept {
        std::swap(h_, o.h_);
        return *this;
    }
    shm_kswitch_key *handle() const { return h_; }

  private:
    shm_kswitch_key *h_ = nullptr;
};

class PhantomGaloisKey {
  public:
    PhantomGaloisKey() = default;
    explicit PhantomGaloisKey(shm_galois_keys *h) : h_(h) {}
    ~PhantomGaloisKey() {
        if (h_) shm_galois_keys_destroy(h_);
    }
    PhantomGaloisKey(const PhantomGaloisKey &) = delete;
    PhantomGaloisKey &operator=(const PhantomGaloisKey &) = delete;
    PhantomGaloisKey(PhantomGaloisKey &&o) noexcept : h_(o.h_) { o.h_ = nullptr; }
    PhantomGaloisKey &operator=(PhantomGaloisKey &&o) noexcept {
        std::swap(h_, o.h_);
        return *this;
    }
    shm_galois_keys *handle() const { return h_; }

  private:
    shm_galois_keys *h_ = nullptr;
};

class PhantomPublicKey {
  public:
    PhantomPublicKey() = default;
    explicit PhantomPublicKey(shm_public_key *h) : h_(h) {}
    ~PhantomPublicKey() {
        if (h_) shm_public_key_destroy(h_);
    }
    PhantomPublicKey(const PhantomPublicKey &) = delete;
    PhantomPublicKey &operator=(const PhantomPublicKey &) = delete;
    PhantomPublicKey(PhantomPublicKey &&o) noexcept : h_(o.h_) { o.h_ = nullptr; }
    PhantomPublicKey &operator=(PhantomPublicKey &&o) noexcept {
        std::swap(h_, o.h_);
        return *this;
    }
    shm_public_key *handle() const { return h_; }

  private:
    shm_public_key *h_ = nullptr;
};

// PhantomSecretKey (secretkey.h:18-309) with injected randomness.
class PhantomSecretKey {
  public:
    PhantomSecretKey(const PhantomContext &ctx, uint64_t seed, uint32_t hamming_weight) : ctx_(&ctx) {
        shm_check(shm_secret_key_generate(ctx.handle(), seed, hamming_weight, &h_));
    }
    ~PhantomSecretKey() {
        if (h_) shm_secret_key_destroy(h_);
    }
    PhantomSecretKey(const PhantomSecretKey &) = delete;
    PhantomSecretKey &operator=(const PhantomSecretKey &) = delete;
    shm_secret_key *handle() const { return h_; }
    PhantomPublicKey gen_publickey(const PhantomContext &ctx, uint64_t seed) const {
        shm_public_key *pk = nullptr;
        shm_check(shm_public_key_generate(ctx.handle(), h_, seed, &pk));
        return PhantomPublicKey(pk);
    }
    PhantomRelinKey gen_relinkey(const PhantomContext &ctx, uint64_t seed) const {
        shm_kswitch_key *k = nullptr;
        shm_check(shm_relin_key_generate(ctx.handle(), h_, seed, &k));
        return PhantomRelinKey(k);
    }
    PhantomGaloisKey create_galois_keys_from_steps(const PhantomContext &ctx, const std::vector<int> &steps,
                                                   uint64_t seed) const {
        shm_galois_keys *g = nullptr;
        shm_check(shm_galois_keys_generate(ctx.handle(), h_, steps.data(), steps.size(), seed, &g));
        return PhantomGaloisKey(g);
    }
    void decrypt(const PhantomContext &ctx, const PhantomCiphertext &ct, PhantomPlaintext &pt) const {
        pt.ensure(ctx, (uint32_t)ct.chain_index());
        shm_check(shm_decrypt(ctx.handle(), h_, ct.handle(), pt.handle(), nullptr));
    }

  private:
    shm_secret_key *h_ = nullptr;
    const PhantomContext *ctx_ = nullptr;
};

// PhantomCKKSEncoder (ckks.h): full slots, N/2.
class PhantomCKKSEncoder {
  public:
    explicit PhantomCKKSEncoder(const PhantomContext &ctx) : ctx_(&ctx) {}
    size_t slot_count() const { return ctx_->poly_modulus_degree() / 2; }
    const PhantomContext &context() const { return *ctx_; }

  private:
    const PhantomContext *ctx_;
};

// moai::Encoder (ckks_evaluator_parallel.cuh:13-105)
class Encoder {
  public:
    Encoder() = default;
    Encoder(PhantomContext *context, PhantomCKKSEncoder *encoder) : context_(context), encoder_(encoder) {}
    size_t slot_count() const { return encoder_->slot_count(); }
    void reset_sparse_slots() {}  // full slots only
    void encode(std::vector<double> values, size_t chain_index, double scale, PhantomPlaintext &plain) {
        if (values.size() == 1) return encode(values[0], chain_index, scale, plain);
        values.resize(slot_count(), 0.0);
        plain.ensure(*context_, (uint32_t)chain_index);
        shm_check(shm_encode(context_->handle(), values.data(), nullptr, values.size(), (uint32_t)chain_index, scale,
                             plain.handle(), nullptr));
    }
    void encode(std::vector<double> values, double scale, PhantomPlaintext &plain) {
        encode(std::move(values), 1, scale, plain);
    }
    void encode(std::vector<std::complex<double>> values, double scale, PhantomPlaintext &plain) {
        if (values.size() == 1) return encode(values[0], scale, plain);
        values.resize(slot_count(), 0.0);
        std::vector<double> re(values.size()), im(values.size());
        for (size_t i = 0; i < values.size(); i++) {
            re[i] = values[i].real();
            im[i] = values[i].imag();
        }
        plain.ensure(*context_, 1);
        shm_check(shm_encode(context_->handle(), re.data(), im.data(), re.size(), 1, scale, plain.handle(), nullptr));
    }
    // a value fills every slot
    void encode(double value, size_t chain_index, double scale, PhantomPlaintext &plain) {
        encode(std::vector<double>(slot_count(), value), chain_index, scale, plain);
    }
    void encode(double value, double scale, PhantomPlaintext &plain) { encode(value, 1, scale, plain); }
    void encode(std::complex<double> value, double scale, PhantomPlaintext &plain) {
        encode(std::vector<std::complex<double>>(slot_count(), value), scale, plain);
    }
    void decode(PhantomPlaintext &plain, std::vector<double> &values) {
        values.assign(slot_count(), 0.0);
        shm_check(shm_decode(context_->handle(), plain.handle(), values.data(), nullptr, values.size(), nullptr));
    }
    void decode(PhantomPlaintext &plain, std::vector<std::complex<double>> &values) {
        std::vector<double> re(slot_count()), im(slot_count());
        shm_check(shm_decode(context_->handle(), plain.handle(), re.data(), im.data(), re.size(), nullptr));
        values.resize(re.size());
        for (size_t i = 0; i < re.size(); i++) values[i] = {re[i], im[i]};
    }

  private:
    PhantomContext *context_ = nullptr;
    PhantomCKKSEncoder *encoder_ = nullptr;
};

// moai::Encryptor (ckks_evaluator_parallel.cuh:107-127): public-key encryption; call
// k of this object draws its randomness from seed + k.
class Encryptor {
  public:
    Encryptor() = default;
    Encryptor(PhantomContext *context, PhantomPublicKey *pk, uint64_t seed = 0) : context_(context), pk_(pk), seed_(seed) {}
    void encrypt(PhantomPlaintext &plain, PhantomCiphertext &ct) {
        if (!ct.handle()) ct = PhantomCiphertext(*context_, 2, 1);
        shm_check(shm_encrypt_asymmetric(context_->handle(), pk_->handle(), plain.handle(), seed_ + calls_++,
                                         ct.handle(), nullptr));
    }

  private:
    PhantomContext *context_ = nullptr;
    PhantomPublicKey *pk_ = nullptr;
    uint64_t seed_ = 0, calls_ = 0;
};

// moai::Decryptor (ckks_evaluator_parallel.cuh:564-598)
class Decryptor {
  public:
    Decryptor() = default;
    Decryptor(PhantomContext *context, PhantomSecretKey *sk) : context_(context), sk_(sk) {}
    void decrypt(PhantomCiphertext &ct, PhantomPlaintext &plain) { sk_->decrypt(*context_, ct, plain); }
    void create_galois_keys_from_steps(std::vector<int> &steps, PhantomGaloisKey &galois_keys, uint64_t seed = 0) {
        galois_keys = sk_->create_galois_keys_from_steps(*context_, steps, seed);
    }

  private:
    PhantomContext *context_ = nullptr;
    PhantomSecretKey *sk_ = nullptr;
};

// moai::Evaluator (ckks_evaluator_parallel.cuh:150-562)
class Evaluator {
  public:
    Evaluator() = default;
    Evaluator(PhantomContext *context, PhantomCKKSEncoder *encoder) : context_(context), encoder_(encoder) {}

    void mod_switch_to_next_inplace(PhantomCiphertext &ct) { shm_check(shm_mod_switch_to_next_inplace(c(), ct.handle(), nullptr)); }
    void mod_switch_to_inplace(PhantomCiphertext &ct, size_t chain_index) {
        shm_check(shm_mod_switch_to_inplace(c(), ct.handle(), (uint32_t)chain_index, nullptr));
    }
    void mod_switch_to_inplace(PhantomPlaintext &pt, size_t chain_index) {
        shm_check(shm_plain_mod_switch_to_inplace(c(), pt.handle(), (uint32_t)chain_index, nullptr));
    }
    void rescale_to_next_inplace(PhantomCiphertext &ct) { shm_check(shm_rescale_to_next_inplace(c(), ct.handle(), nullptr)); }
    void relinearize_inplace(PhantomCiphertext &ct, const PhantomRelinKey &relin_keys) {
        shm_check(shm_relinearize_inplace(c(), ct.handle(), relin_keys.handle(), nullptr));
    }
    void square(PhantomCiphertext &ct, PhantomCiphertext &dest) { multiply(ct, ct, dest); }
    void square_inplace(PhantomCiphertext &ct) { shm_check(shm_square_inplace(c(), ct.handle(), nullptr)); }
    void multiply(PhantomCiphertext &ct1, const PhantomCiphertext &ct2, PhantomCiphertext &dest) {
        if (&ct2 == &dest) {
            multiply_inplace(dest, ct1);
        } else {
            dest = ct1;
            multiply_inplace(dest, &ct1 == &ct2 ? dest : ct2);
        }
    }
    void multiply_inplace(PhantomCiphertext &ct1, const PhantomCiphertext &ct2) {
        shm_check(shm_multiply_inplace(c(), ct1.handle(), ct2.handle(), nullptr));
    }
    void multiply_plain(PhantomCiphertext &ct, PhantomPlaintext &plain, PhantomCiphertext &dest) {
        dest = ct;
        multiply_plain_inplace(dest, plain);
    }
    void multiply_plain_inplace(PhantomCiphertext &ct, PhantomPlaintext &plain) {
        shm_check(shm_multiply_plain_inplace(c(), ct.handle(), plain.handle(), nullptr));
    }
    void add_plain(const PhantomCiphertext &ct, PhantomPlaintext &plain, PhantomCiphertext &dest) {
        dest = ct;
        add_plain_inplace(dest, plain);
    }
    void add_plain_inplace(PhantomCiphertext &ct, PhantomPlaintext &plain) {
        shm_check(shm_add_plain_inplace(c(), ct.handle(), plain.handle(), nullptr));
    }
    void add(PhantomCiphertext &ct1, const PhantomCiphertext &ct2, PhantomCiphertext &dest) {
        if (&ct2 == &dest) {
            add_inplace(dest, ct1);
        } else {
            dest = ct1;
            add_inplace(dest, &ct1 == &ct2 ? dest : ct2);
        }
    }
    void add_inplace(PhantomCiphertext &ct1, const PhantomCiphertext &ct2) {
        shm_check(shm_add_inplace(c(), ct1.handle(), ct2.handle(), nullptr));
    }
    void add_many(std::vector<PhantomCiphertext> &cts, PhantomCiphertext &dest) {
        if (cts.size() < 2) throw std::invalid_argument("add_many requires at least 2 ciphertexts");
        add(cts[0], cts[1], dest);
        for (size_t i = 2; i < cts.size(); i++) add_inplace(dest, cts[i]);
    }
    void sub_plain(PhantomCiphertext &ct, PhantomPlaintext &plain, PhantomCiphertext &dest) {
        dest = ct;
        sub_plain_inplace(dest, plain);
    }
    void sub_plain_inplace(PhantomCiphertext &ct, PhantomPlaintext &plain) {
        shm_check(shm_sub_plain_inplace(c(), ct.handle(), plain.handle(), nullptr));
    }
    void sub(PhantomCiphertext &ct1, const PhantomCiphertext &ct2, PhantomCiphertext &dest) {
        if (&ct2 == &dest) {
            PhantomCiphertext t(ct1);
            sub_inplace(t, ct2);
            dest = std::move(t);
        } else {
            dest = ct1;
            sub_inplace(dest, &ct1 == &ct2 ? dest : ct2);
        }
    }
    void sub_inplace(PhantomCiphertext &ct1, const PhantomCiphertext &ct2) {
        shm_check(shm_sub_inplace(c(), ct1.handle(), ct2.handle(), nullptr));
    }
    void rotate_vector(PhantomCiphertext &ct, int steps, PhantomGaloisKey &galois_keys, PhantomCiphertext &dest) {
        dest = ct;
        rotate_vector_inplace(dest, steps, galois_keys);
    }
    void rotate_vector_inplace(PhantomCiphertext &ct, int steps, PhantomGaloisKey &galois_keys) {
        shm_check(shm_rotate_vector_inplace(c(), ct.handle(), steps, galois_keys.handle(), nullptr));
    }
    void negate(PhantomCiphertext &ct, PhantomCiphertext &dest) {
        dest = ct;
        negate_inplace(dest);
    }
    void negate_inplace(PhantomCiphertext &ct) { shm_check(shm_negate_inplace(c(), ct.handle(), nullptr)); }
    void apply_galois(PhantomCiphertext &ct, uint32_t elt, PhantomGaloisKey &galois_keys, PhantomCiphertext &dest) {
        dest = ct;
        shm_check(shm_apply_galois_inplace(c(), dest.handle(), elt, galois_keys.handle(), nullptr));
    }
    void apply_galois_inplace(PhantomCiphertext &ct, int step, PhantomGaloisKey &galois_keys) {
        shm_check(shm_apply_galois_inplace(c(), ct.handle(), context_->get_elt_from_step(step), galois_keys.handle(),
                                           nullptr));
    }
    void complex_conjugate(PhantomCiphertext &ct, const PhantomGaloisKey &galois_keys, PhantomCiphertext &dest) {
        dest = ct;
        complex_conjugate_inplace(dest, galois_keys);
    }
    void complex_conjugate_inplace(PhantomCiphertext &ct, const PhantomGaloisKey &galois_keys) {
        shm_check(shm_complex_conjugate_inplace(c(), ct.handle(), galois_keys.handle(), nullptr));
    }
    // transform_from_ntt / transform_to_ntt (ckks_evaluator_parallel.cuh:396-442): every
    // polynomial of the ciphertext over its limbs
    void transform_from_ntt_inplace(PhantomCiphertext &ct) { ntt(ct, false); }
    void transform_to_ntt_inplace(PhantomCiphertext &ct) { ntt(ct, true); }
    void transform_from_ntt(const PhantomCiphertext &ct, PhantomCiphertext &dest) {
        dest = ct;
        transform_from_ntt_inplace(dest);
    }
    void transform_to_ntt(const PhantomCiphertext &ct, PhantomCiphertext &dest) {
        dest = ct;
        transform_to_ntt_inplace(dest);
    }
    void multiply_const(const PhantomCiphertext &ct, double value, PhantomCiphertext &dest) {
        dest = ct;
        multiply_const_inplace(dest, value);
    }
    void multiply_const_inplace(PhantomCiphertext &ct, double value) {
        shm_check(shm_multiply_const_inplace(c(), ct.handle(), value, nullptr));
    }
    void add_const(PhantomCiphertext &ct, double value, PhantomCiphertext &dest) {
        dest = ct;
        add_const_inplace(dest, value);
    }
    void add_const_inplace(PhantomCiphertext &ct, double value) {
        shm_check(shm_add_const_inplace(c(), ct.handle(), value, nullptr));
    }
    void add_reduced_error(const PhantomCiphertext &ct1, const PhantomCiphertext &ct2, PhantomCiphertext &dest) {
        if (&ct2 == &dest) {
            add_inplace_reduced_error(dest, ct1);
        } else {
            dest = ct1;
            add_inplace_reduced_error(dest, ct2);
        }
    }
    void add_inplace_reduced_error(PhantomCiphertext &ct1, const PhantomCiphertext &ct2) {
        shm_check(shm_add_inplace_reduced_error(c(), ct1.handle(), ct2.handle(), nullptr));
    }
    void sub_reduced_error(const PhantomCiphertext &ct1, const PhantomCiphertext &ct2, PhantomCiphertext &dest) {
        dest = ct1;
        sub_inplace_reduced_error(dest, ct2);
    }
    void sub_inplace_reduced_error(PhantomCiphertext &ct1, const PhantomCiphertext &ct2) {
        shm_check(shm_sub_inplace_reduced_error(c(), ct1.handle(), ct2.handle(), nullptr));
    }
    void multiply_reduced_error(const PhantomCiphertext &ct1, const PhantomCiphertext &ct2,
                                const PhantomRelinKey &relin_keys, PhantomCiphertext &dest) {
        if (&ct2 == &dest) {
            multiply_inplace_reduced_error(dest, ct1, relin_keys);
        } else {
            dest = ct1;
            multiply_inplace_reduced_error(dest, ct2, relin_keys);
        }
    }
    void multiply_inplace_reduced_error(PhantomCiphertext &ct1, const PhantomCiphertext &ct2,
                                        const PhantomRelinKey &relin_keys) {
        shm_check(shm_multiply_inplace_reduced_error(c(), ct1.handle(), ct2.handle(), relin_keys.handle(), nullptr));
    }
    void double_inplace(PhantomCiphertext &ct) const { shm_check(shm_double_inplace(c(), ct.handle(), nullptr)); }
    void multiply_vector_reduced_error(PhantomCiphertext &ct, std::vector<double> &values, PhantomCiphertext &dest) {
        dest = ct;
        multiply_vector_inplace_reduced_error(dest, values);
    }
    void multiply_vector_inplace_reduced_error(PhantomCiphertext &ct, std::vector<double> &values) {
        values.resize(encoder_->slot_count(), 0.0);
        shm_check(shm_multiply_vector_inplace_reduced_error(c(), ct.handle(), values.data(), nullptr, values.size(),
                                                            nullptr));
    }
    void multiply_vector_inplace_reduced_error(PhantomCiphertext &ct, std::vector<std::complex<double>> &values) {
        values.resize(encoder_->slot_count(), 0.0);
        std::vector<double> re(values.size()), im(values.size());
        for (size_t i = 0; i < values.size(); i++) {
            re[i] = values[i].real();
            im[i] = values[i].imag();
        }
        shm_check(shm_multiply_vector_inplace_reduced_error(c(), ct.handle(), re.data(), im.data(), re.size(), nullptr));
    }

  private:
    shm_context *c() const { return context_->handle(); }
    void ntt(PhantomCiphertext &ct, bool forward) {
        const uint32_t npolys = (uint32_t)ct.size(), nl = (uint32_t)ct.coeff_modulus_size();
        uint64_t *data = shm_ciphertext_data(ct.handle());
        shm_check(forward ? shm_ntt_forward(c(), data, npolys, nl, 0, nullptr)
                          : shm_ntt_inverse(c(), data, npolys, nl, 0, nullptr));
    }
    PhantomContext *context_ = nullptr;
    PhantomCKKSEncoder *encoder_ = nullptr;
};

// ct_pt_matrix_mul_wo_pre (Ct_pt_matrix_mul.cuh:18-185): output i = rescale(sum_j
// enc_X[j] * encode(W[j][i])), j < row_W, i < col_W; W is row_W x col_W.
inline std::vector<PhantomCiphertext> ct_pt_matrix_mul_wo_pre(std::vector<PhantomCiphertext> &enc_X,
                                                              const std::vector<std::vector<double>> &W, int col_X,
                                                              int col_W, int row_W, PhantomContext &context) {
    (void)col_X;
    if (enc_X.empty() || W.empty() || col_W <= 0 || row_W <= 0)
        throw std::invalid_argument("ct_pt_matrix_mul_wo_pre: empty inputs or bad dimensions");
    if ((int)enc_X.size() < row_W || (int)W.size() < row_W)
        throw std::invalid_argument("ct_pt_matrix_mul_wo_pre: fewer inputs or weight rows than row_W");
    std::vector<double> w((size_t)row_W * col_W);
    for (int j = 0; j < row_W; j++) {
        if ((int)W[j].size() < col_W) throw std::invalid_argument("ct_pt_matrix_mul_wo_pre: short weight row");
        for (int i = 0; i < col_W; i++) w[(size_t)j * col_W + i] = W[j][i];
    }
    std::vector<const shm_ciphertext *> xs(row_W);
    for (int j = 0; j < row_W; j++) xs[j] = enc_X[j].handle();
    std::vector<PhantomCiphertext> out;
    out.reserve(col_W);
    for (int i = 0; i < col_W; i++) out.emplace_back(context, 2, 1);
    std::vector<shm_ciphertext *> os(col_W);
    for (int i = 0; i < col_W; i++) os[i] = out[i].handle();
    shm_check(shm_ct_pt_matrix_mul(context.handle(), xs.data(), (uint32_t)row_W, w.data(), (uint32_t)col_W, os.data(),
                                   SHM_MATMUL_INT8, nullptr));
    return out;
}

// batch_input (Batch_encode_encrypt.cuh:12-43): X[j][k][i] -> num_col ciphertexts
inline std::vector<PhantomCiphertext> batch_input(const std::vector<std::vector<std::vector<double>>> &X, int num_X,
                                                  int num_row, int num_col, double scale, PhantomContext &context,
                                                  PhantomPublicKey &pk, uint64_t seed = 0) {
    std::vector<double> flat((size_t)num_X * num_row * num_col);
    for (int j = 0; j < num_X; j++)
        for (int k = 0; k < num_row; k++)
            for (int i = 0; i < num_col; i++) flat[((size_t)j * num_row + k) * num_col + i] = X.at(j).at(k).at(i);
    std::vector<PhantomCiphertext> out;
    out.reserve(num_col);
    for (int i = 0; i < num_col; i++) out.emplace_back(context, 2, 1);
    std::vector<shm_ciphertext *> os(num_col);
    for (int i = 0; i < num_col; i++) os[i] = out[i].handle();
    shm_check(shm_batch_input(context.handle(), pk.handle(), flat.data(), num_X, num_row, num_col, scale, seed,
                              os.data(), nullptr));
    return out;
}

// bias_vec (Batch_encode_encrypt.cuh:45-57)
inline std::vector<int> bias_vec(const std::vector<int> &lengths, int num_X, int num_row) {
    std::vector<int> out((size_t)num_X * num_row);
    shm_check(shm_bias_vec(lengths.data(), num_X, num_row, out.data(), out.size()));
    return out;
}

// The bootstrapping plan and its keys (Bootstrapper.cuh:63; the reference's CKKSEvaluator
// carries the relinearisation and Galois keys the batch bootstrapper uses).
class Bootstrapper {
  public:
    Bootstrapper(PhantomContext &ctx, const shm_boot_params &params) : ctx_(&ctx) {
        shm_check(shm_bootstrapper_create(ctx.handle(), &params, &h_));
    }
    ~Bootstrapper() {
        if (h_) shm_bootstrapper_destroy(h_);
    }
    Bootstrapper(const Bootstrapper &) = delete;
    Bootstrapper &operator=(const Bootstrapper &) = delete;
    std::vector<int> rotation_steps() const {
        size_t n = 0;
        shm_check(shm_bootstrapper_rotation_steps(h_, nullptr, 0, &n));
        std::vector<int> v(n);
        shm_check(shm_bootstrapper_rotation_steps(h_, v.data(), n, &n));
        return v;
    }
    shm_bootstrapper *handle() const { return h_; }
    PhantomContext &context() const { return *ctx_; }

  private:
    PhantomContext *ctx_;
    shm_bootstrapper *h_ = nullptr;
};

struct CKKSEvaluator {
    PhantomContext *context = nullptr;
    PhantomRelinKey *relin_keys = nullptr;
    PhantomGaloisKey *galois_keys = nullptr;
};

// BatchBootstrapper (BatchBootstrap.cuh:470-820), amortised mode (keys resident).
class BatchBootstrapper {
  public:
    BatchBootstrapper(Bootstrapper &boot, CKKSEvaluator &ckks, size_t n) : boot_(boot), ckks_(&ckks), batch_n_(n) {
        if (n == 0) throw std::invalid_argument("BatchBootstrapper: batch size must be positive");
    }
    ~BatchBootstrapper() {
        if (keys_) shm_shared_a_keys_destroy(keys_);
    }
    BatchBootstrapper(const BatchBootstrapper &) = delete;
    BatchBootstrapper &operator=(const BatchBootstrapper &) = delete;
    void prepare() { shm_check(shm_bootstrapper_prepare(boot_.handle(), nullptr)); }
    // one-time shared-a key material at the inputs' level, plus the shared rotation keys
    void prepare_shared_keys(PhantomSecretKey &secret_key, size_t input_chain_index, uint32_t hamming_weight = 192,
                             uint64_t seed = 0) {
        if (keys_) shm_shared_a_keys_destroy(keys_);
        keys_ = nullptr;
        shm_check(shm_shared_a_keys_create(boot_.context().handle(), secret_key.handle(), (uint32_t)batch_n_,
                                           hamming_weight, seed, (uint32_t)input_chain_index, &keys_));
        const std::vector<int> steps = boot_.rotation_steps();
        shm_check(shm_shared_a_keys_add_galois(keys_, steps.data(), steps.size(), seed + 1));
    }
    void enable_chunk_key_amortization(size_t) {}  // keys always stay resident
    void bootstrap_ordinary_s2c_first(std::vector<PhantomCiphertext> &rtn, std::vector<PhantomCiphertext> &cts,
                                      double &keygen_sec, double &transform_sec) {
        (void)keygen_sec;
        (void)transform_sec;
        run(rtn, cts, false);
    }
    void bootstrap_shared_a_s2c_first(std::vector<PhantomCiphertext> &rtn, std::vector<PhantomCiphertext> &cts,
                                      double &keygen_sec, double &transform_sec, double layer3_factor = 1.0 / 16.0) {
        (void)keygen_sec;
        (void)transform_sec;
        (void)layer3_factor;  // fixed by the bootstrapper's parameters
        if (!keys_) throw std::logic_error("BatchBootstrapper: call prepare_shared_keys first");
        if (cts.size() != batch_n_)
            throw std::invalid_argument("BatchBootstrapper: shared-a bootstrap needs exactly batch_n ciphertexts");
        run(rtn, cts, true);
    }

  private:
    void run(std::vector<PhantomCiphertext> &rtn, std::vector<PhantomCiphertext> &cts, bool shared) {
        std::vector<const shm_ciphertext *> in(cts.size());
        for (size_t i = 0; i < cts.size(); i++) in[i] = cts[i].handle();
        rtn.clear();
        for (size_t i = 0; i < cts.size(); i++) rtn.emplace_back(boot_.context(), 2, 1);
        std::vector<shm_ciphertext *> out(cts.size());
        for (size_t i = 0; i < cts.size(); i++) out[i] = rtn[i].handle();
        if (shared)
            shm_check(shm_bootstrap_shared_a_s2c_first(boot_.handle(), keys_, in.data(), in.size(), out.data(),
                                                       ckks_->relin_keys->handle(), 0.0, nullptr));
        else
            shm_check(shm_bootstrap_s2c_first(boot_.handle(), in.data(), in.size(), out.data(),
                                              ckks_->galois_keys->handle(), ckks_->relin_keys->handle(), 0.0,
                                              nullptr));
    }
    Bootstrapper &boot_;
    CKKSEvaluator *ckks_;
    size_t batch_n_;
    shm_shared_a_keys *keys_ = nullptr;
};

}  // namespace moai

#endif /* SHMOAI_HPP */
