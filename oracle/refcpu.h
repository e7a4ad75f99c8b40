/*
 * refcpu — CPU restatement of ShMOAI/PhantomFHE limb arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP engine in
 * shmoai_amd/: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it. The product path never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the ShMOAI repository root; Phantom sources live under
 * thirdparty/phantom-fhe/).
 *
 * Conventions (identical to the reference):
 *   - a polynomial at level Ql is uint64 limb-major [limb][N];
 *   - limb k of a "QlP" polynomial uses prime Q[k] for k < Ql and the special
 *     prime P[k-Ql] (global prime index nQ + k - Ql) otherwise;
 *   - ciphertexts are [poly][limb][N] in NTT (bit-reversed evaluation) form;
 *   - a key-switching key is dnum digits, each [2][QP][N] (b-part, a-part).
 */
#ifndef SHMOAI_REFCPU_H
#define SHMOAI_REFCPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct RefCtx RefCtx;

/* context ---------------------------------------------------------------- */
RefCtx *ref_ctx_new(int logN, int n_bits, const int *bit_sizes, int n_special);
void ref_ctx_free(RefCtx *c);
size_t ref_ctx_nq(const RefCtx *c);
size_t ref_ctx_np(const RefCtx *c);
void ref_ctx_primes(const RefCtx *c, uint64_t *out);           /* nQ+nP */
void ref_ctx_psi(const RefCtx *c, uint64_t *out);              /* nQ+nP */
void ref_ctx_tables(const RefCtx *c, size_t prime, uint64_t *tw, uint64_t *itw);
void ref_set_threads(int n);

/* number theory helpers exposed for pinning ------------------------------- */
int ref_is_prime(uint64_t v);
void ref_get_primes(size_t N, int bit_size, size_t count, uint64_t *out);

/* NTT ---------------------------------------------------------------------- */
void ref_ntt(const RefCtx *c, uint64_t *data, size_t nlimbs, const uint32_t *prime_idx);
void ref_intt(const RefCtx *c, uint64_t *data, size_t nlimbs, const uint32_t *prime_idx);

/* elementwise (Ql data limbs, npolys polynomials) ------------------------- */
void ref_add(const RefCtx *c, const uint64_t *a, const uint64_t *b, uint64_t *out, size_t Ql, size_t npolys);
void ref_sub(const RefCtx *c, const uint64_t *a, const uint64_t *b, uint64_t *out, size_t Ql, size_t npolys);
void ref_negate(const RefCtx *c, const uint64_t *a, uint64_t *out, size_t Ql, size_t npolys);
void ref_multiply_plain(const RefCtx *c, uint64_t *ct, const uint64_t *pt, size_t Ql, size_t npolys);
void ref_tensor(const RefCtx *c, const uint64_t *a, const uint64_t *b, uint64_t *out3, size_t Ql);

/* level management -------------------------------------------------------- */
void ref_rescale(const RefCtx *c, const uint64_t *in, uint64_t *out, size_t Ql, size_t npolys);
void ref_mod_switch_drop(const RefCtx *c, const uint64_t *in, uint64_t *out, size_t Ql, size_t npolys);

/* Galois ------------------------------------------------------------------ */
uint32_t ref_galois_elt_from_step(const RefCtx *c, int step);
void ref_galois_table(const RefCtx *c, uint32_t elt, uint32_t *table);
void ref_apply_galois_ntt(const RefCtx *c, uint32_t elt, const uint64_t *in, uint64_t *out, size_t nlimbs);

/* hybrid key switching ------------------------------------------------------ */
size_t ref_beta(const RefCtx *c, size_t Ql);
void ref_modup(const RefCtx *c, const uint64_t *c2, size_t Ql, uint64_t *out);
void ref_ks_inner(const RefCtx *c, const uint64_t *modup, const uint64_t *key, size_t Ql, uint64_t *cx);
void ref_moddown(const RefCtx *c, uint64_t *cx, size_t Ql, uint64_t *out);
void ref_keyswitch(const RefCtx *c, const uint64_t *c2, const uint64_t *key, size_t Ql, uint64_t *ks0, uint64_t *ks1);
void ref_rotate(const RefCtx *c, const uint64_t *ct, size_t Ql, uint32_t elt, const uint64_t *key, uint64_t *out);
void ref_relinearize(const RefCtx *c, const uint64_t *ct3, size_t Ql, const uint64_t *key, uint64_t *out);

/* samplers + keys + encryption --------------------------------------------- */
uint64_t ref_splitmix64(uint64_t key, uint64_t ctr);
void ref_sample_uniform(const RefCtx *c, uint64_t seed, size_t nlimbs, const uint32_t *prime_idx, uint64_t *out);
void ref_sample_gaussian(const RefCtx *c, uint64_t seed, int64_t *out);
void ref_sample_ternary(const RefCtx *c, uint64_t seed, size_t hamming_weight, int64_t *out);
void ref_sk_ntt(const RefCtx *c, const int64_t *coeffs, uint64_t *sk_qp);
void ref_gen_kswitch_key(const RefCtx *c, const uint64_t *sk_qp, const uint64_t *newkey_qp, uint64_t seed, uint64_t *key);
void ref_gen_relin_key(const RefCtx *c, const uint64_t *sk_qp, uint64_t seed, uint64_t *key);
void ref_gen_galois_key(const RefCtx *c, const uint64_t *sk_qp, uint32_t elt, uint64_t seed, uint64_t *key);
void ref_encrypt_symmetric(const RefCtx *c, const uint64_t *sk_qp, const uint64_t *pt, size_t Ql, uint64_t seed, uint64_t *ct);
void ref_decrypt(const RefCtx *c, const uint64_t *sk_qp, const uint64_t *ct, size_t size, size_t Ql, uint64_t *pt);
void ref_sample_ternary_uniform(const RefCtx *c, uint64_t seed, int64_t *out);
void ref_gen_public_key(const RefCtx *c, const uint64_t *sk_qp, uint64_t seed, uint64_t *pk);
void ref_encrypt_asymmetric(const RefCtx *c, const uint64_t *pk, const uint64_t *pt, uint64_t seed, uint64_t *ct);

/* CKKS encoder (FP64) ------------------------------------------------------- */
void ref_encode(const RefCtx *c, const double *re, const double *im, size_t n_values, size_t sparse_slots,
                size_t Ql, double scale, uint64_t *pt);
void ref_encode_scalar(const RefCtx *c, double value, size_t Ql, double scale, uint64_t *pt);
void ref_decode(const RefCtx *c, const uint64_t *pt, size_t Ql, size_t sparse_slots, double scale,
                double *re, double *im);

/* MOAI ct-pt matmul --------------------------------------------------------- */
void ref_modraise(const RefCtx *c, const uint64_t *in, uint64_t *out);
void ref_mul_monomial(const RefCtx *c, uint64_t *data, size_t nlimbs, size_t npolys, size_t power);
void ref_extend_exact(const RefCtx *c, const uint64_t *coef, size_t Ql, uint64_t *outP);
void ref_gen_fmt_key(const RefCtx *c, const uint64_t *sk_qp, const uint64_t *const *aux_qp, size_t n, size_t Ql,
                     uint64_t seed, uint64_t *a_fmt, uint64_t *B);
void ref_fmt_switch_forward(const RefCtx *c, const uint64_t *const *cts, size_t n, size_t Ql, const uint64_t *a_fmt,
                            const uint64_t *B, uint64_t *out);
void ref_gen_shared_galois_key(const RefCtx *c, const uint64_t *const *aux_qp, size_t n, uint32_t elt, uint64_t seed,
                               uint64_t *A, uint64_t *B);
void ref_ctpt_matmul(const RefCtx *c, const uint64_t *const *X, size_t row_w, const double *W, size_t col_w,
                     size_t Ql, const double *scales, size_t col_begin, size_t col_end, uint64_t *const *out);

#ifdef __cplusplus
}
#endif
#endif
