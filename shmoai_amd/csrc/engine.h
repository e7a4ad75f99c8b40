// Host-side engine structures shared by the HIP translation units.
//
// Layouts (HBM):
//   polynomial at level Ql        uint64 [Ql][N]              (limb-major)
//   ciphertext                     uint64 [size][Ql][N]       NTT form
//   key-switching key              uint64 [dnum][2][QP][N]    b-part, a-part
//   twiddles per prime             uint64 [QP][N]             bit-reversed powers
// Prime index convention: 0..nQ-1 data primes, nQ..nQ+nP-1 special primes.
// Limb k of a "QlP" polynomial uses prime k for k < Ql and nQ + (k - Ql) after.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "arith.h"

namespace shm {

// Error classes surfaced through the C ABI (mirrors the reference's exception
// types: std::invalid_argument, std::logic_error, std::runtime_error).
struct InvalidArgument : std::invalid_argument {
    using std::invalid_argument::invalid_argument;
};
struct LogicError : std::logic_error {
    using std::logic_error::logic_error;
};
struct RuntimeError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char *what);
#define SHM_HIP(call) ::shm::hip_check((call), #call)

// Stream-ordered device blocks: hipMallocAsync on stream s, returned with
// hipFreeAsync, except that blocks of the default stream are cached and reused
// on the default stream (context.hip).
void *pool_alloc(size_t bytes, hipStream_t s, int *dev_out);
void pool_free(void *p, size_t bytes, hipStream_t s, int dev);

// Device buffer owned by the host object (freed with the object). alloc() is a
// plain hipMalloc (freed after draining the device); alloc_async() takes the
// block from the stream-ordered pool and returns it on the last stream the
// buffer was used on, so no host synchronisation is needed.
class DevBuf {
  public:
    DevBuf() = default;
    explicit DevBuf(size_t bytes) { alloc(bytes); }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : ptr_(o.ptr_), bytes_(o.bytes_), s_(o.s_), async_(o.async_), dev_(o.dev_) {
        o.ptr_ = nullptr;
        o.bytes_ = 0;
    }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            release();
            ptr_ = o.ptr_;
            bytes_ = o.bytes_;
            s_ = o.s_;
            async_ = o.async_;
            dev_ = o.dev_;
            o.ptr_ = nullptr;
            o.bytes_ = 0;
        }
        return *this;
    }
    void alloc(size_t bytes);
    void alloc_async(size_t bytes, hipStream_t s);
    // order the eventual free after the work already queued on s
    void use_on(hipStream_t s);
    void release();
    template <class T>
    T *as() const {
        return reinterpret_cast<T *>(ptr_);
    }
    size_t bytes() const { return bytes_; }

  private:
    void *ptr_ = nullptr;
    size_t bytes_ = 0;
    hipStream_t s_ = nullptr;
    bool async_ = false;
    int dev_ = 0;
};

// Stream-ordered scratch on the op's stream (pool_alloc / pool_free).
class Scratch {
  public:
    Scratch(size_t bytes, hipStream_t s);
    ~Scratch();
    Scratch(const Scratch &) = delete;
    Scratch &operator=(const Scratch &) = delete;
    template <class T>
    T *as() const {
        return reinterpret_cast<T *>(ptr_);
    }

  private:
    void *ptr_ = nullptr;
    size_t bytes_ = 0;
    hipStream_t s_;
    int dev_ = 0;
};

// Per-level constants (level = number of data limbs Ql).
struct LevelConsts {
    size_t Ql = 0, beta = 0;
    // rescale: q_last^-1 mod q_i, i < Ql-1 (value, shoup)
    DevBuf qlast_inv, qlast_inv_s;
    // modup: digit-local hat inverse per data limb i (value, shoup)
    DevBuf hat_inv, hat_inv_s;
    // modup: [qhat_i]_{p_k} for digit g, local i, out limb k: [beta][alpha][QlP]
    DevBuf hat_mod;
    // modup: [D_g]_{p_k}, D_g = product of digit g's primes: [beta][QlP]
    DevBuf dig_mod;
    // ModDown merged with the rescale: basis B = {q_last, p_0 .. p_(nP-1)} (the
    // limbs Ql-1 .. Ql+nP-1 of a QlP polynomial). bhat_inv[j] = (B/b_j)^-1 mod b_j
    // (value, shoup); bhat_mod[i][j] = (B/b_j) mod q_i, i < Ql-1; bmod[i] = B mod
    // q_i; binv[i] = B^-1 mod q_i (value, shoup)
    DevBuf bhat_inv, bhat_inv_s, bhat_mod, bmod, binv, binv_s;
};

struct Context {
    int logN = 0;
    size_t N = 0, nQ = 0, nP = 0, dnum = 0;
    std::vector<uint64_t> q;  // QP primes
    std::vector<uint64_t> psi;
    int device = 0;

    // device tables
    DevBuf d_mod;           // DevMod[QP]
    DevBuf tw, tw_s;        // [QP][N]
    DevBuf itw, itw_s;      // [QP][N]
    DevBuf ninv, ninv_s;    // [QP]
    DevBuf pinv, pinv_s;    // P^-1 mod q_i, i < nQ
    DevBuf pmod, pmod_s;    // P mod q_i, i < nQ
    DevBuf phat_inv, phat_inv_s;  // [nP] (P/p_j)^-1 mod p_j
    DevBuf phat_mod;        // [nQ][nP] (P/p_j) mod q_i
    std::vector<std::unique_ptr<LevelConsts>> levels;  // index Ql (0 unused)

    // psi^(N/2) per prime (value, shoup): the NTT-slot factor of X^(N/2)
    DevBuf iota, iota_s;

    // encoder
    DevBuf enc_roots;       // complex<double>[2N] (re, im interleaved)
    DevBuf enc_group;       // uint32[N/4]

    // Galois permutation tables cache: elt -> uint32[N]
    std::mutex galois_mu;
    std::map<uint32_t, std::unique_ptr<DevBuf>> galois_tables;

    size_t nQP() const { return nQ + nP; }
    // every prime in (2^45, 2^52): base conversions take the 26-bit split products
    bool primes52 = false;
    size_t ql_of_chain(size_t chain) const { return nQ - chain + 1; }
    size_t chain_of_ql(size_t Ql) const { return nQ - Ql + 1; }
    const DevMod *mods() const { return d_mod.as<DevMod>(); }
    const LevelConsts &level(size_t Ql) const { return *levels.at(Ql); }
    const uint32_t *galois_table(uint32_t elt, hipStream_t s);
    uint32_t elt_from_step(int step) const;

    // A second stream of the context (non-blocking, default priority), created on
    // first use, for work that overlaps the caller's stream inside one call.
    struct AuxStream {
        std::mutex mu;
        hipStream_t s = nullptr;
        ~AuxStream();
    } aux;
    hipStream_t aux_stream();
};

std::unique_ptr<Context> make_context(int logN, const std::vector<int> &bits, size_t nP);

struct Ciphertext {
    Ciphertext() = default;
    Ciphertext(Ciphertext &&) = default;
    Ciphertext &operator=(Ciphertext &&) = default;
    Context *ctx = nullptr;
    size_t size = 0;      // polynomials (2 or 3)
    size_t chain = 1;     // chain index (1 = all of Q)
    size_t cap_limbs = 0; // allocated limbs per polynomial
    size_t cap_size = 0;
    double scale = 1.0;
    bool ntt = true;
    DevBuf buf;
    size_t Ql() const { return ctx->ql_of_chain(chain); }
    uint64_t *data() const { return buf.as<uint64_t>(); }
    uint64_t *poly(size_t i) const { return data() + i * Ql() * ctx->N; }
    // (re)allocates from the stream-ordered pool on s when the capacity is too small;
    // contents are kept only when no reallocation happens
    void resize(size_t new_size, size_t new_chain, hipStream_t s = nullptr, bool zero = true);
};

struct Plaintext {
    Context *ctx = nullptr;
    size_t chain = 1;
    double scale = 1.0;
    DevBuf buf;
    size_t Ql() const { return ctx->ql_of_chain(chain); }
    uint64_t *data() const { return buf.as<uint64_t>(); }
};

struct SecretKey {
    Context *ctx = nullptr;
    DevBuf s;  // NTT form over QP: [QP][N]
};

// secretkey.h PhantomPublicKey: (-(a s + e), a) over QP in NTT form, [2][QP][N]
struct PublicKey {
    Context *ctx = nullptr;
    DevBuf pk;
};

struct KSwitchKey {
    Context *ctx = nullptr;
    DevBuf k;  // [dnum][2][QP][N]
};

struct GaloisKeys {
    Context *ctx = nullptr;
    std::map<uint32_t, std::unique_ptr<KSwitchKey>> keys;
};

// ---- kernels / launchers (defined in the .hip files) ----------------------
// NTT over n_limbs consecutive limbs whose primes are given by prime_idx
// (device array) or by a contiguous range starting at first_prime.
void ntt_forward(const Context &c, uint64_t *data, size_t n_polys, size_t n_limbs, size_t first_prime,
                 hipStream_t s);
// Events recorded around the tile pass of every fused-rescale NTT launched while
// set (nullptr to stop); used to time the matmul's rescale.
void set_rescale_tile_events(hipEvent_t start, hipEvent_t stop);
// Timing probe of the fused last NTT pass (ModDown combine and rescale) over eager
// launches: begin clears and arms it; end waits for the device and returns the
// launch count, the summed kernel time and the algorithmic bytes of those launches.
void probe_fused_ntt_begin();
void probe_fused_ntt_end(size_t *launches, double *ms, double *bytes);
// The same window's plain forward tile passes (the ModUp digits' and other forward
// NTTs' last pass), with 16 N algorithmic bytes per limb; valid after probe_fused_ntt_end.
void probe_plain_ntt_last(size_t *launches, double *ms, double *bytes);
// Key-switch work counters, per level Ql (index < 128): polynomials ModUp'd,
// inner-product output polynomials, polynomials ModDown'd since the last reset.
void ks_counters(uint64_t *modup, uint64_t *inner, uint64_t *moddown, size_t cap);
void ks_counters_reset();
// Launch-shape knobs of the key-switch inner products (results do not depend on
// them): log2 of the member slots per block of the batched inner product and of
// the output slots of the multi-output one (0..4; -1: the autotuned choice).
struct KsTuning {
    int batch_slots_log2 = -1;
    int multi_slots_log2 = -1;
    // the INT8 ct-pt matmul (row_W <= 256) overlaps its contraction with the rescale on
    // the context's second stream (1) or runs them one after the other (0, the default:
    // both kernels fill the GPU, and the overlapped form measured 7% slower)
    int matmul_overlap = 0;
    // threads per block of the NTT tile passes: 512 (radix-8 rounds, more waves per
    // 32 KB tile; the default) or 256 (radix-16 rounds); the results are the same words
    int ntt_tile_threads = 512;
};
KsTuning &ks_tuning();
// Autotuning of those slot counts on first use of each shape (on by default): switch
// it on or off, forget the cached choices, or list them with their timings.
void ks_autotune_enable(bool on);
void ks_autotune_clear();
std::string ks_autotune_report();
// Encoder decompose + forward NTT in one: out[p][k] = NTT of the rounded FFT
// output fft[p] (N/2 complex entries: real parts are coefficients 0..N/2-1,
// imaginary parts N/2..N-1) reduced mod prime k, for k < n_limbs.
void ntt_forward_encoded(const Context &c, const double2 *fft, uint64_t *out, size_t n_polys, size_t n_limbs,
                         hipStream_t s);
void ntt_inverse(const Context &c, uint64_t *data, size_t n_polys, size_t n_limbs, size_t first_prime,
                 hipStream_t s);
// Same, for a QlP-shaped polynomial: limbs [0, Ql) use primes 0..Ql-1 and the
// rest use special primes; only limbs [lo, hi) are transformed, skipping
// [skip_lo, skip_hi).
void ntt_forward_qlp(const Context &c, uint64_t *data, size_t Ql, size_t lo, size_t hi, size_t skip_lo,
                     size_t skip_hi, size_t n_polys, size_t poly_stride, hipStream_t s);
void ntt_inverse_qlp(const Context &c, uint64_t *data, size_t Ql, size_t lo, size_t hi, size_t n_polys,
                     size_t poly_stride, hipStream_t s);

// Rescale tail fused into one forward NTT: limb k (< L1) of polynomial p is
// NTT([last_p]_{q_k}); the block pass stores (acc[p][k] - that) * qinv_k to
// outs[p / ppo] (or out_single) at ((p % ppo) * L1 + k) * N.
// rounded: divide with rounding to nearest (centred last limb) instead of the floor.
// rescale_ntt_fused over output limbs lo .. lo + n - 1 of L1 only (outputs keep L1 limbs per polynomial)
void rescale_ntt_fused_window(const Context &c, const uint64_t *last, const uint64_t *acc, size_t acc_limbs,
                              uint64_t *delta, uint64_t *const *outs, size_t ppo, size_t n_polys, size_t L1,
                              size_t lo, size_t n, const uint64_t *qinv, const uint64_t *qinv_s, hipStream_t s);
void rescale_ntt_fused(const Context &c, const uint64_t *last, const uint64_t *acc, size_t acc_limbs,
                       uint64_t *delta, uint64_t *const *outs, uint64_t *out_single, size_t ppo, size_t n_polys,
                       size_t L1, const uint64_t *qinv, const uint64_t *qinv_s, hipStream_t s, bool rounded = false);

// Per-polynomial base pointers for batched calls over polynomials that live in
// separate buffers (passed to kernels by value).
constexpr int kPolyPack = 64;
struct PolyPack {
    uint64_t *p[kPolyPack];
};
// Forward NTT of the ModDown correction delta (n_polys x Ql limbs, in place)
// whose last pass writes (cx - NTT(delta)) * pinv to out.p[p] (adding onto it
// when bit p of acc_mask is set); cx holds cx_limbs limbs per polynomial.
void moddown_ntt_fused(const Context &c, uint64_t *delta, const uint64_t *cx, size_t Ql, size_t cx_limbs,
                       const PolyPack &out, size_t n_polys, uint64_t acc_mask, const uint64_t *pinv,
                       const uint64_t *pinv_s, hipStream_t s, const PolyPack *add_src = nullptr,
                       const uint32_t *add_perm = nullptr);
// ntt_inverse of polynomials held in separate buffers: limb k of polynomial p is read
// from in.p[p] + k N (n_limbs limbs each, at most 64 polynomials) and the result is
// written to out ([n_polys][n_limbs][N]); the inputs are left as they are.
void ntt_inverse_gather(const Context &c, const PolyPack &in, uint64_t *out, size_t n_polys, size_t n_limbs,
                        size_t first_prime, hipStream_t s);
// The fused rescale over n_polys polynomials in separate buffers: input p at
// in.p[p] ([L1 + 1][N]), output p at out.p[p] ([L1][N]).
void rescale_ntt_fused_pack(const Context &c, const uint64_t *last, const PolyPack &in, const PolyPack &out,
                            uint64_t *delta, size_t n_polys, size_t L1, const uint64_t *qinv, const uint64_t *qinv_s,
                            hipStream_t s, bool rounded);

// elementwise (poly.hip)
void launch_add(const Context &c, uint64_t *a, const uint64_t *b, size_t Ql, size_t npolys, hipStream_t s);
void launch_sub(const Context &c, uint64_t *a, const uint64_t *b, size_t Ql, size_t npolys, hipStream_t s);
void launch_negate(const Context &c, uint64_t *a, size_t Ql, size_t npolys, hipStream_t s);
void launch_mul_plain(const Context &c, uint64_t *ct, const uint64_t *pt, size_t Ql, size_t npolys, hipStream_t s);
// operands with more limbs than Ql (a_limbs / b_limbs per polynomial, 0 = Ql) are
// used through their first Ql limbs
void launch_tensor(const Context &c, const uint64_t *a, const uint64_t *b, uint64_t *out3, size_t Ql, hipStream_t s,
                   size_t a_limbs = 0, size_t b_limbs = 0, bool accumulate = false);
void launch_square(const Context &c, const uint64_t *a, uint64_t *out3, size_t Ql, hipStream_t s,
                   size_t a_limbs = 0);
// (d0, d1, d2) = 2 (a (x) b) - cval (t0, t1, 0), the Chebyshev step T_k = 2 T_a T_b - T_(a-b)
// before relinearisation (b == a squares; t may be null); operands through their first Ql limbs
void launch_tensor_cheb(const Context &c, const uint64_t *a, const uint64_t *b, const uint64_t *t, uint64_t cval,
                        uint64_t *out3, size_t Ql, hipStream_t s, size_t a_limbs, size_t b_limbs, size_t t_limbs);
void launch_rescale(const Context &c, const uint64_t *in, uint64_t *out, size_t Ql, size_t npolys, hipStream_t s,
                    bool rounded = false);
// the same with caller-provided scratch: last [npolys][N], delta [npolys][Ql - 1][N]
void launch_rescale_scratch(const Context &c, const uint64_t *in, uint64_t *out, size_t Ql, size_t npolys,
                            hipStream_t s, bool rounded, uint64_t *last, uint64_t *delta);
void launch_permute(const Context &c, const uint64_t *in, uint64_t *out, const uint32_t *table, size_t nlimbs,
                    hipStream_t s);
// out[p][i] = in.p[p][i], i < words, p < npolys
void launch_gather_polys(const Context &c, const PolyPack &in, uint64_t *out, size_t words, size_t npolys,
                         hipStream_t s);
// poly 0 (+)= [rounded]_q per limb; npolys polynomials (*)= [rounded]_q per limb
// (rounded = round(value * scale), as encode_scalar computes it)
void launch_add_scalar(const Context &c, uint64_t *a, double rounded, size_t Ql, hipStream_t s);
// Batched elementwise ops over separate polynomials of Ql limbs: op 0 adds the
// rounded constant, op 1 multiplies by it, op 2 adds b.p[i] onto a.p[i].
void launch_pack_op(const Context &c, const PolyPack &a, const PolyPack &b, size_t npolys, double rounded, size_t Ql,
                    int op, hipStream_t s);
void launch_mul_scalar(const Context &c, uint64_t *a, double rounded, size_t Ql, size_t npolys, hipStream_t s);
// zero fill by a kernel (a memset node recorded into a captured graph is avoided)
void launch_fill_zero(uint64_t *out, size_t words, hipStream_t s);
// device-to-device copy by a kernel (non-overlapping ranges)
void launch_copy_words(uint64_t *dst, const uint64_t *src, size_t words, hipStream_t s);
void launch_copy_limbs(const Context &c, const uint64_t *in, size_t in_limbs, uint64_t *out, size_t out_limbs,
                       size_t npolys, hipStream_t s);
// samplers
void launch_sample_uniform(const Context &c, uint64_t seed, uint64_t *out, size_t Ql, bool withP, hipStream_t s);
void launch_sample_gaussian_rns(const Context &c, uint64_t seed, uint64_t *out, size_t Ql, bool withP,
                                hipStream_t s);
void launch_signed_to_rns(const Context &c, const int8_t *coeffs, uint64_t *out, size_t Ql, bool withP,
                          hipStream_t s);
void launch_ks_key_digit(const Context &c, const uint64_t *sk, const uint64_t *newkey, uint64_t *b,
                         const uint64_t *a, const uint64_t *e, size_t digit, hipStream_t s);
void launch_encrypt_combine(const Context &c, const uint64_t *sk, const uint64_t *a, const uint64_t *e,
                            const uint64_t *pt, uint64_t *c0, size_t Ql, hipStream_t s);
void launch_decrypt(const Context &c, const uint64_t *ct, size_t size, const uint64_t *sk, uint64_t *pt,
                    size_t Ql, hipStream_t s);
void launch_mul_elementwise(const Context &c, const uint64_t *a, const uint64_t *b, uint64_t *out, size_t nlimbs,
                            bool withP, size_t Ql, hipStream_t s);
// uniform ternary coefficients (splitmix64(seed, j) mod 3) - 1 as residues over Ql (+P) limbs
void launch_sample_ternary_uniform(const Context &c, uint64_t seed, uint64_t *out, size_t Ql, bool withP,
                                   hipStream_t s);
// over all QP limbs: b = -(a s + e)
void launch_neg_as_plus_e(const Context &c, const uint64_t *sk, const uint64_t *a, const uint64_t *e, uint64_t *b,
                          hipStream_t s);
// over all QP limbs, npolys polynomials: out[p] = u * pk[p] + e
void launch_pk_mac(const Context &c, const uint64_t *u, const uint64_t *pk, const uint64_t *e, uint64_t *out,
                   size_t npolys, hipStream_t s);

// key switching (keyswitch.hip)
// rounded: ModDown rounds to nearest with an exact base conversion (unbiased) instead
// of the reference's floor with the approximate conversion.
void keyswitch(const Context &c, const uint64_t *c2, const uint64_t *key, size_t Ql, uint64_t *ct0, uint64_t *ct1,
               bool add_to_ct1, hipStream_t s, bool rounded = false);

// key-switching pieces (keyswitch.hip)
constexpr int kMaxKeys = 34;
constexpr int kMaxDigits = 16;
struct KeyRefs {
    const uint64_t *base[kMaxKeys];  // output o reads digit g at base[o] + g * stride[o]
    size_t stride[kMaxKeys];
};
// ModUp of one Ql-limb NTT polynomial into [beta][QlP][N] (NTT form); rounded =
// exact centred digit lift instead of the reference's approximate conversion
void ks_modup(const Context &c, const uint64_t *c2, size_t Ql, uint64_t *d, hipStream_t s, bool rounded = false);
// cx[o] = sum_{g < ndig} d_g * key_o(g) over QlP limbs, o < nout; with perm, the
// digits are read through that NTT-domain permutation (hoisted rotations)
void ks_inner_multi(const Context &c, const uint64_t *d, size_t Ql, size_t ndig, const KeyRefs &kr, size_t nout,
                    bool local, uint64_t *cx, hipStream_t s, const uint32_t *perm = nullptr, bool accumulate = false);
// ModDown of npolys [QlP][N] polynomials into out + p * out_stride (bit p of acc_mask: accumulate)
void ks_moddown(const Context &c, uint64_t *cx, size_t npolys, size_t Ql, uint64_t *out, size_t out_stride,
                uint64_t acc_mask, hipStream_t s, bool rounded = false);
// The same with polynomial p written to out.p[p]. With add_src and perm, a polynomial p
// whose acc_mask bit is set gets add_src.p[p] read through the NTT-domain permutation
// perm added instead of out.p[p] (a Galois permutation fused into the accumulation).
void ks_moddown_pack(const Context &c, uint64_t *cx, size_t npolys, size_t Ql, const PolyPack &out,
                     uint64_t acc_mask, hipStream_t s, bool rounded = false, const PolyPack *add_src = nullptr,
                     const uint32_t *perm = nullptr);
// ModUp of m polynomials c2.p[i] (Ql limbs, NTT) into d [m][beta][QlP][N]; from 4
// members on the digits' own limbs are not written (they are c2's limbs: pass c2
// as `own` to ks_inner_batch, which reads them from there in every case)
void ks_modup_batch(const Context &c, const PolyPack &c2, size_t m, size_t Ql, uint64_t *d, hipStream_t s,
                    bool rounded = false);
// cx [m][2][QlP][N] = inner products of the m digit sets of d with one key
// ([dnum][2][QP][N]); each key word is read once for the group; with perm the
// digits are read through that Galois permutation; with `add` (2m polynomials of
// Ql limbs, NTT form), P * add.p[2i + o] is added onto the data limbs of cx[i][o]
void ks_inner_batch(const Context &c, const uint64_t *d, size_t m, size_t Ql, const uint64_t *key, uint64_t *cx,
                    hipStream_t s, const uint32_t *perm = nullptr, const PolyPack *own = nullptr,
                    const PolyPack *add = nullptr);
// Relinearisation merged with the following rescale (inside bootstrapping): for m
// products with tensor parts d01.p[2i], d01.p[2i + 1] and c2.p[i] (Ql limbs, NTT
// form), out.p[2i + o] (Ql - 1 limbs) = round((P d_o + ks_o(c2)) / (P q_last)), one
// exact centred conversion over {q_last, P} in place of ModDown then rescale.
void keyswitch_rescale_batch(const Context &c, const PolyPack &c2, const PolyPack &d01, size_t m, const uint64_t *key,
                             size_t Ql, const PolyPack &out, hipStream_t s);
// Key switch of m polynomials c2.p[i] with one key; the switched pair of member
// i goes to (ct01.p[2i], ct01.p[2i + 1]), accumulated where acc_mask has the
// polynomial's bit. One launch per stage for the whole group.
void keyswitch_batch(const Context &c, const PolyPack &c2, size_t m, const uint64_t *key, size_t Ql,
                     const PolyPack &ct01, uint64_t acc_mask, hipStream_t s, bool rounded = false);

// encoder (encoder.hip)
void encode(const Context &c, const double *re, const double *im, size_t n_values, size_t Ql, double scale,
            uint64_t *out, hipStream_t s);
void encode_scalar(const Context &c, double value, size_t Ql, double scale, uint64_t *out, hipStream_t s);
// encode() of the full slot vector {re + i im, ..., re + i im} (device-filled, no host data)
void encode_const(const Context &c, double re, double im, size_t Ql, double scale, uint64_t *out, hipStream_t s);
// B plaintexts [B][Ql][N]: plaintext b encodes column col0 + b of the host-layout
// array X [num_x][num_row][num_col] (device memory) in batch_input's interleaved
// packing, slot num_x * k + j = X[j][k][col0 + b]; fix[b] = scale / (N/2) (device)
void encode_columns_batch(const Context &c, const double *X, size_t num_x, size_t num_row, size_t num_col,
                          size_t col0, size_t B, size_t Ql, const double *fix, uint64_t *out, double2 *fftbuf,
                          hipStream_t s);
// B plaintexts [B][Ql][N] in NTT form: plaintext b encodes the slot vector
// {mask[t] ? w[b] : 0} (t < N/2) at scale fix[b] * N/2, exactly as encode() does
// for that vector. w, fix, mask are device arrays; fftbuf holds B * N/2 double2.
void encode_masked_batch(const Context &c, const double *w, const double *fix, const uint8_t *mask, size_t B,
                         size_t Ql, uint64_t *out, double2 *fftbuf, hipStream_t s);
void decode(const Context &c, const uint64_t *pt, size_t Ql, double scale, double *re, double *im, size_t n,
            hipStream_t s);

// MOAI matmul (matmul.hip)
// Limb windows of the INT8 contraction: the last limb is contracted first and
// last_done() runs right after it is queued, then the other limbs in windows of
// `window`, window_done(lo, n) after each (both on the caller's thread, in stream order).
struct Int8WindowHook {
    size_t window = 8;
    std::function<void()> last_done;
    std::function<void(size_t, size_t)> window_done;
};
void ctpt_matmul(const Context &c, const std::vector<const Ciphertext *> &X, const double *W, size_t row_w,
                 size_t col_w, const std::vector<Ciphertext *> &out, int variant, hipStream_t s);
float last_matmul_rescale_tile_ms(); // tile pass of the last batched matmul's fused rescale NTT
// masked form: the plaintext of term (j, i) encodes {mask[t] ? W[j][i] : 0} (t < N/2)
void ctpt_matmul_masked(const Context &c, const std::vector<const Ciphertext *> &X, const double *W, size_t row_w,
                        size_t col_w, const uint8_t *mask, const std::vector<Ciphertext *> &out, hipStream_t s);

// evaluator-level operations (ops.hip)
uint64_t host_splitmix64(uint64_t key, uint64_t ctr);
std::vector<int8_t> ternary_secret(size_t N, uint64_t seed, uint32_t hw);
void load_secret(Context &c, SecretKey &sk, const int8_t *coeffs);
void gen_kswitch(Context &c, const SecretKey &sk, const uint64_t *newkey_dev, uint64_t seed, KSwitchKey &out);
// `rounded` selects the unbiased (round-to-nearest) rescale / ModDown used inside
// bootstrapping; the default is the reference's floor semantics.
void apply_galois(Context &c, Ciphertext &ct, uint32_t elt, const GaloisKeys &gk, hipStream_t s, bool rounded = false);
void rotate_internal(Context &c, Ciphertext &ct, int step, const GaloisKeys &gk, hipStream_t s, bool rounded = false);
void rescale_inplace(Context &c, Ciphertext &ct, hipStream_t s, bool rounded = false);
// rescale_inplace with its scratch in the library's static device array (noalloc.hip):
// no device allocation; calls that overlap on one GPU race on that array
void rescale_inplace_noalloc(Context &c, Ciphertext &ct, hipStream_t s);
size_t noalloc_scratch_words();
// rescale_inplace of ciphertexts at one level in one pass (<= 64 polynomials)
void rescale_batch(Context &c, const std::vector<Ciphertext *> &cts, hipStream_t s, bool rounded = false);
// Rotations of one ciphertext by several steps sharing one ModUp of c1 (hoisting,
// evaluate.cu:1727-1920 hoisting_inplace); outs[i] receives rot(ct, steps[i]).
// Steps without their own key fall back to rotate_internal.
void rotate_hoisted(Context &c, const Ciphertext &ct, const std::vector<int> &steps, const GaloisKeys &gk,
                    const std::vector<Ciphertext *> &outs, hipStream_t s, bool rounded = false);
// evaluate.cu:1727-1920 hoisting_inplace: ct <- sum_i rot(ct, steps[i]) with one
// ModUp of c1 and one ModDown of the accumulated inner products
void hoisting_inplace(Context &c, Ciphertext &ct, const std::vector<int> &steps, const GaloisKeys &gk,
                      hipStream_t s);
// The same operations on a group of ordinary ciphertexts held as one container
// of 2m polynomials (b_0, a_0, b_1, a_1, ...), m <= 32, batched over the members
void apply_galois_group(Context &c, Ciphertext &g, uint32_t elt, const GaloisKeys &gk, hipStream_t s,
                        bool rounded = false);
void rotate_internal_group(Context &c, Ciphertext &g, int step, const GaloisKeys &gk, hipStream_t s,
                           bool rounded = false);
void rotate_hoisted_group(Context &c, const Ciphertext &g, const std::vector<int> &steps, const GaloisKeys &gk,
                          const std::vector<Ciphertext *> &outs, hipStream_t s, bool rounded = false);
void mod_switch_drop(Context &c, Ciphertext &ct, size_t new_chain, hipStream_t s);
// public-key encryption (encrypt.hip)
void gen_public_key(Context &c, const SecretKey &sk, uint64_t seed, PublicKey &out, hipStream_t s);
void encrypt_asymmetric(Context &c, const PublicKey &pk, const uint64_t *pt, size_t pt_chain, double pt_scale,
                        uint64_t seed, Ciphertext &out, hipStream_t s);
// batch_input: X host [num_x][num_row][num_col]; out[i] = Enc_pk(encode(column i))
void batch_input(Context &c, const PublicKey &pk, const double *X, size_t num_x, size_t num_row, size_t num_col,
                 double scale, uint64_t seed, const std::vector<Ciphertext *> &out, hipStream_t s);
// facade operations of moai::Evaluator (facade.hip)
void facade_multiply_const(Context &c, Ciphertext &ct, double re, double im, hipStream_t s);
void facade_add_const(Context &c, Ciphertext &ct, double re, double im, hipStream_t s);
void facade_multiply_vector(Context &c, Ciphertext &ct, const double *re, const double *im, size_t n,
                            hipStream_t s);
// op: 0 add, 1 sub, 2 multiply (+ relinearize with rk)
void facade_reduced_error(Context &c, Ciphertext &ct1, const Ciphertext &ct2, int op, const KSwitchKey *rk,
                          hipStream_t s);
void ct_clone(Ciphertext &dst, const Ciphertext &src, hipStream_t s);
void ct_add(Context &c, Ciphertext &a, const Ciphertext &b, hipStream_t s);
void ct_sub(Context &c, Ciphertext &a, const Ciphertext &b, hipStream_t s);
void ct_add_const(Context &c, Ciphertext &a, double v, hipStream_t s);
// ct_add_const / ct_mul_const / ct_add over ciphertexts of one level (and one scale
// for the constant forms) in one launch each; each member gets exactly the words of
// the single-ciphertext operation
void ct_add_const_batch(Context &c, const std::vector<Ciphertext *> &a, double v, hipStream_t s);
void ct_mul_const_batch(Context &c, const std::vector<Ciphertext *> &a, double v, double sigma, hipStream_t s);
void ct_add_batch(Context &c, const std::vector<Ciphertext *> &a, const std::vector<const Ciphertext *> &b,
                  hipStream_t s);
void ct_mul_const_rescale(Context &c, Ciphertext &a, double v, double sigma, hipStream_t s, bool rounded = false);
void ct_align(Context &c, Ciphertext &a, size_t chain, double scale, hipStream_t s, bool rounded = false);
void ct_mul_relin_rescale(Context &c, Ciphertext &out, const Ciphertext &a, const Ciphertext &b,
                          const KSwitchKey &rk, hipStream_t s, bool rounded = false);
// ct_mul_relin_rescale of out[i] = a[i] * b[i], i < m <= 32, batched
// With `cheb` (rounded only), each product is the Chebyshev step 2 a b - cheb[i]
// (cheb[i] may be null: 2 a b), cheb[i] entering at the product's level and scale
// through an integer constant folded into the tensor before the relinearisation.
void mul_relin_rescale_batch(Context &c, const std::vector<Ciphertext *> &out,
                             const std::vector<const Ciphertext *> &a, const std::vector<const Ciphertext *> &b,
                             const KSwitchKey &rk, hipStream_t s, bool rounded = false,
                             const std::vector<const Ciphertext *> *cheb = nullptr);
// Ct_ct_matrix_mul.cuh:10-181 ct_ct_matrix_mul_colpacking: out[i] =
// rescale(relinearize(sum_j X[j] * rot(W[j], i * num_batch))) for i < out.size().
void ctct_matmul_colpacking(Context &c, const std::vector<const Ciphertext *> &X,
                            const std::vector<const Ciphertext *> &W, int num_batch, const GaloisKeys &gk,
                            const KSwitchKey &rk, const std::vector<Ciphertext *> &out, hipStream_t s);
// Ct_ct_matrix_mul.cuh:182-609 ct_ct_matrix_mul_diagpacking (baby-step giant-step);
// X holds the col_X ciphertexts the product reads, out one per W ciphertext.
void ctct_matmul_diagpacking(Context &c, const std::vector<const Ciphertext *> &X,
                             const std::vector<const Ciphertext *> &W, int num_batch, const GaloisKeys &gk,
                             const KSwitchKey &rk, const std::vector<Ciphertext *> &out, hipStream_t s);
// a <- a * [v encoded at scale sigma] (no rescale)
void ct_mul_const(Context &c, Ciphertext &a, double v, double sigma, hipStream_t s);
void ct_mul_monomial_half(Context &c, Ciphertext &a, bool negate, hipStream_t s);
void ct_modraise(Context &c, Ciphertext &a, hipStream_t s);

// timing hook for the contraction kernel(s) of the last batched matmul (HIP events)
float last_matmul_kernel_ms();
int last_matmul_variant();

}  // namespace shm
