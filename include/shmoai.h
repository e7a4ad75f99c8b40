/*
 * shmoai.h — C ABI of the MI355X-native CKKS engine (libshmoai.so).
 *
 * This is the drop-in boundary for ShMOAI's encrypted-BERT hot path. The
 * reference exposes no C/FFI interface: its API is the C++ `moai::` facade
 * (src/include/source/ckks_evaluator_parallel.cuh:13-680) over PhantomFHE free
 * functions (thirdparty/phantom-fhe/include/evaluate.cuh:32-304) and classes
 * (context.cuh:136, secretkey.h:18-309, ciphertext.h:21, ckks.h). Each entry
 * point below names the reference interface it replaces.
 *
 * Conventions
 *  - Every function returns SHM_OK (0) or a negative SHM_ERR_* code; the
 *    message of the last failure on the calling thread is shm_last_error().
 *    Error classes mirror the reference's exceptions: std::invalid_argument ->
 *    SHM_ERR_INVALID_ARGUMENT, std::logic_error -> SHM_ERR_LOGIC,
 *    std::runtime_error -> SHM_ERR_RUNTIME; the engine never calls exit().
 *  - Streams are hipStream_t passed as void*; NULL is the legacy default
 *    stream. Operations only enqueue work, except where noted (import/export,
 *    decode, key generation synchronise the stream before returning).
 *  - chain_index follows the reference: 1 = all data primes, +1 per dropped
 *    prime (context.cu:151-174).
 *  - Device data layout: a ciphertext is uint64 [size][Ql][N] in NTT form, a
 *    plaintext uint64 [Ql][N]; Ql = size_q - chain_index + 1.
 *  - Handles are owned by the caller and released with the matching
 *    *_destroy. Contexts, keys and plaintexts may be shared read-only between
 *    threads that use different streams.
 */
#ifndef SHMOAI_H
#define SHMOAI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SHM_OK 0
#define SHM_ERR_INVALID_ARGUMENT (-1)
#define SHM_ERR_LOGIC (-2)
#define SHM_ERR_RUNTIME (-3)
#define SHM_ERR_OUT_OF_MEMORY (-4)
#define SHM_ERR_DEVICE (-5)

/* ct-pt matmul variants (shm_ct_pt_matrix_mul `variant`) */
#define SHM_MATMUL_U64 0    /* batched 64-bit contraction + batched rescale */
#define SHM_MATMUL_PER_CT 1 /* reference structure: multiply_plain + add per term */
#define SHM_MATMUL_INT8 2   /* batched INT8-MFMA contraction (bit-identical) */

typedef struct shm_context shm_context;
typedef struct shm_ciphertext shm_ciphertext;
typedef struct shm_plaintext shm_plaintext;
typedef struct shm_secret_key shm_secret_key;
typedef struct shm_kswitch_key shm_kswitch_key;
typedef struct shm_galois_keys shm_galois_keys;
typedef void *shm_stream_t;

typedef struct shm_context_info {
    uint32_t poly_degree;  /* N */
    uint32_t size_q;       /* data primes */
    uint32_t size_p;       /* special primes (alpha) */
    uint32_t dnum;         /* size_q / size_p key digits */
} shm_context_info;

/* ---- errors / version ---------------------------------------------------- */
const char *shm_last_error(void);
const char *shm_version(void);

/* Waits until all work queued on `stream` (NULL: the default stream) has
 * finished. Most entry points return with their kernels still queued; the
 * reference's callers use cudaStreamSynchronize on the stream they passed
 * (e.g. ct_pt_matrix_mul_wo_pre, Ct_pt_matrix_mul.cuh:172). */
int shm_synchronize(shm_stream_t stream);

/* Streams and graph capture. The reference runs on phantom::util::cuda_stream_wrapper
 * pool streams (Ct_pt_matrix_mul.cuh:66-106); shm_stream_create makes such a stream
 * (default priority, non-blocking). A sequence of enqueue-only calls on a created
 * stream can be recorded into a graph and replayed: run it once eagerly on that stream
 * first (first uses build device tables with a synchronous upload and move buffers
 * onto the stream), keep the ciphertexts it touches at their sizes (an output that
 * would grow allocates), then shm_graph_capture_begin, the calls,
 * shm_graph_capture_end, and shm_graph_launch as often as needed. Temporaries are
 * allocated and freed inside the graph. Entry points documented as synchronising
 * (matmul, encode/decode with host buffers, import/export, key generation) cannot be
 * captured. A ciphertext or plaintext last used on a caller-owned stream returns its
 * memory on that stream when destroyed: destroy it before that stream. Streams from
 * shm_stream_create are kept alive by the library after shm_stream_destroy (drained and
 * reused), so they are always safe. No counterpart in the reference for the graph calls. */
int shm_stream_create(shm_stream_t *stream);
int shm_stream_destroy(shm_stream_t stream);
typedef struct shm_graph shm_graph;
int shm_graph_capture_begin(shm_stream_t stream);
int shm_graph_capture_end(shm_stream_t stream, shm_graph **graph);
int shm_graph_launch(shm_graph *graph, shm_stream_t stream);
int shm_graph_destroy(shm_graph *graph);

/* ---- context --------------------------------------------------------------
 * Replaces EncryptionParameters + CoeffModulus::Create (modulus.cu:79-108) +
 * PhantomContext(parms) (context.cu:125-238). coeff_bit_sizes lists the data
 * primes followed by the special primes, exactly as passed to
 * CoeffModulus::Create; special_modulus_size = set_special_modulus_size. */
int shm_context_create(uint32_t log_poly_degree, const int *coeff_bit_sizes, uint32_t n_coeff,
                       uint32_t special_modulus_size, shm_context **out);
int shm_context_destroy(shm_context *ctx);
int shm_context_get_info(const shm_context *ctx, shm_context_info *info);
/* all primes, data then special (size_q + size_p words) */
int shm_context_coeff_modulus(const shm_context *ctx, uint64_t *out, size_t n);
/* galois.cu:47-90 PhantomGaloisTool::get_elt_from_step */
int shm_galois_elt_from_step(const shm_context *ctx, int step, uint32_t *elt);

/* ---- raw limb transforms ------------------------------------------------------
 * Replace nwt_2d_radix8_forward_inplace / nwt_2d_radix8_backward_inplace
 * (fntt_2d.cu:620, intt_2d.cu:724): n_polys x n_limbs limbs, limb k uses prime
 * first_prime + k. `data` is a device pointer aligned to 16 bytes (the tile pass
 * reads pairs of words; SHM_ERR_INVALID_ARGUMENT otherwise, before any launch). */
int shm_ntt_forward(shm_context *ctx, uint64_t *data, uint32_t n_polys, uint32_t n_limbs, uint32_t first_prime,
                    shm_stream_t stream);
int shm_ntt_inverse(shm_context *ctx, uint64_t *data, uint32_t n_polys, uint32_t n_limbs, uint32_t first_prime,
                    shm_stream_t stream);

/* ---- ciphertexts / plaintexts (PhantomCiphertext ciphertext.h:21-288,
 * PhantomPlaintext plaintext.h) -------------------------------------------- */
int shm_ciphertext_create(shm_context *ctx, uint32_t size, uint32_t chain_index, shm_ciphertext **out);
int shm_ciphertext_destroy(shm_ciphertext *ct);
int shm_ciphertext_info(const shm_ciphertext *ct, uint32_t *size, uint32_t *chain_index, uint32_t *coeff_mod_size,
                        double *scale);
int shm_ciphertext_set_scale(shm_ciphertext *ct, double scale);
uint64_t *shm_ciphertext_data(shm_ciphertext *ct);
int shm_ciphertext_import(shm_ciphertext *ct, const uint64_t *host, size_t n_words, shm_stream_t stream);
int shm_ciphertext_export(const shm_ciphertext *ct, uint64_t *host, size_t n_words, shm_stream_t stream);
int shm_ciphertext_copy(const shm_ciphertext *src, shm_ciphertext *dst, shm_stream_t stream);

int shm_plaintext_create(shm_context *ctx, uint32_t chain_index, shm_plaintext **out);
int shm_plaintext_destroy(shm_plaintext *pt);
int shm_plaintext_info(const shm_plaintext *pt, uint32_t *chain_index, uint32_t *coeff_mod_size, double *scale);
int shm_plaintext_set_scale(shm_plaintext *pt, double scale);
uint64_t *shm_plaintext_data(shm_plaintext *pt);
int shm_plaintext_import(shm_plaintext *pt, const uint64_t *host, size_t n_words, shm_stream_t stream);
int shm_plaintext_export(const shm_plaintext *pt, uint64_t *host, size_t n_words, shm_stream_t stream);

/* ---- keys (PhantomSecretKey secretkey.cu:431-646) -------------------------
 * Key material is injected: a ternary secret from explicit coefficients, or
 * from a seeded exact-Hamming-weight sampler; uniform `a` polynomials and
 * discrete-Gaussian (sigma 3.2) errors come from SplitMix64 counter streams of
 * the given seed, so the same seed reproduces the same keys on any host. */
int shm_secret_key_create(shm_context *ctx, const int8_t *coeffs, shm_secret_key **out);
int shm_secret_key_generate(shm_context *ctx, uint64_t seed, uint32_t hamming_weight, shm_secret_key **out);
int shm_secret_key_destroy(shm_secret_key *sk);
/* NTT-form secret over all QP primes (size_q + size_p limbs) */
int shm_secret_key_export(const shm_secret_key *sk, uint64_t *host, size_t n_words);
/* secretkey.cu:539-565 gen_relinkey */
int shm_relin_key_generate(shm_context *ctx, const shm_secret_key *sk, uint64_t seed, shm_kswitch_key **out);
/* secretkey.cu:567-575 gen_kswitch_key: switch from new_key (NTT form over QP, host) to sk */
int shm_kswitch_key_generate(shm_context *ctx, const shm_secret_key *sk, const uint64_t *new_key_qp, uint64_t seed,
                             shm_kswitch_key **out);
int shm_kswitch_key_destroy(shm_kswitch_key *key);
/* dnum x [2][QP][N] words */
int shm_kswitch_key_export(const shm_kswitch_key *key, uint64_t *host, size_t n_words);
/* secretkey.cu:634-646 create_galois_keys_from_steps; the key of element elt
 * uses seed splitmix64(seed, elt) */
int shm_galois_keys_generate(shm_context *ctx, const shm_secret_key *sk, const int *steps, size_t n_steps,
                             uint64_t seed, shm_galois_keys **out);
int shm_galois_keys_destroy(shm_galois_keys *gk);
int shm_galois_keys_export(const shm_galois_keys *gk, uint32_t elt, uint64_t *host, size_t n_words);

/* ---- encryption (secretkey.cu:260-326 symmetric, 725-766 decrypt) -------- */
int shm_encrypt_symmetric(shm_context *ctx, const shm_secret_key *sk, const shm_plaintext *pt, uint64_t seed,
                          shm_ciphertext *out, shm_stream_t stream);
int shm_decrypt(shm_context *ctx, const shm_secret_key *sk, const shm_ciphertext *ct, shm_plaintext *out,
                shm_stream_t stream);

/* ---- public-key encryption (PhantomPublicKey, secretkey.h; moai::Encryptor
 * ckks_evaluator_parallel.cuh:107-127) ----------------------------------------------
 * Randomness is injected: the key or ciphertext draws from splitmix64(seed, 0)
 * (uniform a / ternary u) and splitmix64(seed, 1) (the error). */
typedef struct shm_public_key shm_public_key;
/* PhantomSecretKey::gen_publickey (secretkey.cu:524-535): encrypt_zero_symmetric at the
 * key level, pk = (-(a s + e), a) over all (size_q + size_p) primes, NTT form.
 * Synchronises the default stream. */
int shm_public_key_generate(shm_context *ctx, const shm_secret_key *sk, uint64_t seed, shm_public_key **out);
int shm_public_key_destroy(shm_public_key *pk);
/* words [2][size_q + size_p][N] */
int shm_public_key_export(const shm_public_key *pk, uint64_t *host, size_t n_words);
/* PhantomPublicKey::encrypt_asymmetric (secretkey.cu:145-217, CKKS branch) with
 * encrypt_zero_asymmetric_internal(_internal) (secretkey.cu:15-143): u uniform ternary
 * and one error e over the key level, (u pk_0 + e, u pk_1 + e) brought to Q by ModDown
 * (rns_bconv.cu:712-761), plaintext added to c_0; scale = the plaintext's. The
 * plaintext must be at chain_index 1 (the reference's lower-level path applies the
 * level-1 ModDown to a ciphertext of another level). */
int shm_encrypt_asymmetric(shm_context *ctx, const shm_public_key *pk, const shm_plaintext *pt, uint64_t seed,
                           shm_ciphertext *out, shm_stream_t stream);
/* batch_input (Batch_encode_encrypt.cuh:12-43): X is [num_X][num_row][num_col] doubles
 * (row-major, host memory); out[i] (i < num_col) = encrypt_asymmetric(encode(v_i, scale))
 * with v_i[num_X * k + j] = X[j][k][i] and zeros up to the slot count; column i draws
 * its randomness from splitmix64(seed, i). num_X * num_row must not exceed N/2.
 * Synchronises the stream (X is read from host memory). */
int shm_batch_input(shm_context *ctx, const shm_public_key *pk, const double *X, int num_X, int num_row, int num_col,
                    double scale, uint64_t seed, shm_ciphertext *const *out, shm_stream_t stream);
/* bias_vec (Batch_encode_encrypt.cuh:45-57), host only: out[j * num_X + i] = 1 for
 * j < lengths[i], else 0; n = num_X * num_row */
int shm_bias_vec(const int *lengths, int num_X, int num_row, int *out, size_t n);

/* ---- encoder (PhantomCKKSEncoder ckks.cu:89-246; moai::Encoder
 * ckks_evaluator_parallel.cuh:34-103) --------------------------------------- */
int shm_encode(shm_context *ctx, const double *re, const double *im, size_t n_values, uint32_t chain_index,
               double scale, shm_plaintext *out, shm_stream_t stream);
/* moai::Encoder::encode(double value, chain_index, scale, pt): all slots = value */
int shm_encode_scalar(shm_context *ctx, double value, uint32_t chain_index, double scale, shm_plaintext *out,
                      shm_stream_t stream);
/* synchronises the stream; n <= N/2 slots written to re/im (im may be NULL) */
int shm_decode(shm_context *ctx, const shm_plaintext *pt, double *re, double *im, size_t n, shm_stream_t stream);

/* ---- moai::Evaluator facade operations (ckks_evaluator_parallel.cuh:444-561;
 * ckks_evaluator_parallel.cu:466-590). A constant or vector is encoded over all
 * N/2 slots at the ciphertext's scale (moai::Encoder), mod-switched to its level. */
/* multiply_const_inplace: ct *= encode(value); scale becomes scale^2 */
int shm_multiply_const_inplace(shm_context *ctx, shm_ciphertext *ct, double value, shm_stream_t stream);
/* the same with a complex constant (encode(complex<double>)) */
int shm_multiply_const_complex_inplace(shm_context *ctx, shm_ciphertext *ct, double re, double im,
                                       shm_stream_t stream);
/* add_const_inplace: ct += encode(value) */
int shm_add_const_inplace(shm_context *ctx, shm_ciphertext *ct, double value, shm_stream_t stream);
/* multiply_vector_inplace_reduced_error: ct *= encode(values padded with zeros) */
int shm_multiply_vector_inplace_reduced_error(shm_context *ctx, shm_ciphertext *ct, const double *re,
                                              const double *im, size_t n, shm_stream_t stream);
/* add/sub/multiply_inplace_reduced_error: operands at different levels are aligned by
 * multiplying the higher one by s_low q_last / s_high^2, rescaling and mod switching;
 * multiply relinearises. */
int shm_add_inplace_reduced_error(shm_context *ctx, shm_ciphertext *ct1, const shm_ciphertext *ct2,
                                  shm_stream_t stream);
int shm_sub_inplace_reduced_error(shm_context *ctx, shm_ciphertext *ct1, const shm_ciphertext *ct2,
                                  shm_stream_t stream);
int shm_multiply_inplace_reduced_error(shm_context *ctx, shm_ciphertext *ct1, const shm_ciphertext *ct2,
                                       const shm_kswitch_key *relin_keys, shm_stream_t stream);
/* double_inplace: ct += ct */
int shm_double_inplace(shm_context *ctx, shm_ciphertext *ct, shm_stream_t stream);

/* ---- evaluator (evaluate.cuh:32-304) ---------------------------------------- */
int shm_add_inplace(shm_context *ctx, shm_ciphertext *a, const shm_ciphertext *b, shm_stream_t stream);
int shm_sub_inplace(shm_context *ctx, shm_ciphertext *a, const shm_ciphertext *b, shm_stream_t stream);
int shm_negate_inplace(shm_context *ctx, shm_ciphertext *a, shm_stream_t stream);
int shm_add_plain_inplace(shm_context *ctx, shm_ciphertext *a, const shm_plaintext *pt, shm_stream_t stream);
int shm_sub_plain_inplace(shm_context *ctx, shm_ciphertext *a, const shm_plaintext *pt, shm_stream_t stream);
/* evaluate.cu:1237-1266 multiply_plain_ntt */
int shm_multiply_plain_inplace(shm_context *ctx, shm_ciphertext *ct, const shm_plaintext *pt, shm_stream_t stream);
/* evaluate.cu:350-402 (size-3 result) */
int shm_multiply_inplace(shm_context *ctx, shm_ciphertext *a, const shm_ciphertext *b, shm_stream_t stream);
int shm_square_inplace(shm_context *ctx, shm_ciphertext *a, shm_stream_t stream);
/* evaluate.cu:1358-1390 */
int shm_relinearize_inplace(shm_context *ctx, shm_ciphertext *ct, const shm_kswitch_key *relin_key,
                            shm_stream_t stream);
/* eval_key_switch.cu:91-178 keyswitch_inplace with an arbitrary key (size-2 ct) */
int shm_apply_kswitch_key_inplace(shm_context *ctx, shm_ciphertext *ct, const shm_kswitch_key *key,
                                  shm_stream_t stream);
/* evaluate.cu:1563-1584 rescale_to_next (floor division by the last prime) */
int shm_rescale_to_next_inplace(shm_context *ctx, shm_ciphertext *ct, shm_stream_t stream);
/* No counterpart in the reference. rescale_to_next_inplace for callers that cannot
 * allocate device memory (code under graph capture, tight serving loops): the same
 * limbs as shm_rescale_to_next_inplace, with the kernels' scratch (size x N words for
 * the last limbs plus size x (Ql - 1) x N for the correction) in a static 64 MiB device
 * array per GPU instead of the memory pool; larger ciphertexts are refused with
 * SHM_ERR_INVALID_ARGUMENT. Every call on a GPU shares that GPU's array: calls that
 * overlap on one GPU (other streams or threads) race. */
int shm_rescale_to_next_inplace_noalloc(shm_context *ctx, shm_ciphertext *ct, shm_stream_t stream);
/* evaluate.cu:1445-1561 mod_switch_to_next / mod_switch_to (drop limbs) */
int shm_mod_switch_to_next_inplace(shm_context *ctx, shm_ciphertext *ct, shm_stream_t stream);
int shm_mod_switch_to_inplace(shm_context *ctx, shm_ciphertext *ct, uint32_t chain_index, shm_stream_t stream);
int shm_plain_mod_switch_to_inplace(shm_context *ctx, shm_plaintext *pt, uint32_t chain_index, shm_stream_t stream);
/* evaluate.cu:1586-1647 apply_galois_inplace */
int shm_apply_galois_inplace(shm_context *ctx, shm_ciphertext *ct, uint32_t galois_elt, const shm_galois_keys *gk,
                             shm_stream_t stream);
/* evaluate.cu:1650-1715 rotate_vector_inplace (NAF split when a key is missing) */
int shm_rotate_vector_inplace(shm_context *ctx, shm_ciphertext *ct, int step, const shm_galois_keys *gk,
                              shm_stream_t stream);
/* evaluate.cu:1727-1920 hoisting_inplace: ct <- sum of the rotations of ct by steps[],
 * one ModUp of c1 and one ModDown of the summed inner products */
int shm_hoisting_inplace(shm_context *ctx, shm_ciphertext *ct, const shm_galois_keys *gk, const int *steps,
                         size_t n_steps, shm_stream_t stream);
/* evaluate.cu:1717-1725 complex_conjugate_inplace */
int shm_complex_conjugate_inplace(shm_context *ctx, shm_ciphertext *ct, const shm_galois_keys *gk,
                                  shm_stream_t stream);

/* ---- MOAI ct-pt matmul ---------------------------------------------------------
 * Replaces ct_pt_matrix_mul_wo_pre(enc_X, W, col_X, col_W, row_W, context)
 * (src/include/source/matrix_mul/Ct_pt_matrix_mul.cuh:18-185):
 * out[i] = rescale(sum_{j<row_W} X[j] (.) encode(W[j][i], X[j].scale)),
 * out[i].scale = X[0].scale. W is row-major row_W x col_W (W[j*col_W + i]).
 * out[] must hold col_W ciphertexts; they are resized as needed.
 * Synchronises the stream. */
int shm_ct_pt_matrix_mul(shm_context *ctx, const shm_ciphertext *const *enc_X, uint32_t row_W, const double *W,
                         uint32_t col_W, shm_ciphertext *const *out, int variant, shm_stream_t stream);
/* Replaces ct_pt_matrix_mul_wo_pre_w_mask / _w_mask_single
 * (Ct_pt_matrix_mul.cuh:846-985, 987-1100): as shm_ct_pt_matrix_mul, but the
 * plaintext of term (j, i) encodes the slot vector {bias_vec[t] == 1 ? W[j][i] : 0}
 * (t < slot_count = N/2) at X[j]'s scale, a full polynomial per term. The terms are
 * encoded on the device in blocks (FP64 special FFT as in shm_encode, then NTT) and
 * contracted with 128-bit sums. n_bias must equal N/2. Synchronises the stream. */
int shm_ct_pt_matrix_mul_masked(shm_context *ctx, const shm_ciphertext *const *enc_X, uint32_t row_W,
                                const double *W, uint32_t col_W, const int32_t *bias_vec, size_t n_bias,
                                shm_ciphertext *const *out, shm_stream_t stream);
/* duration of the dominant kernel of the last matmul on this process, ms (HIP events) */
double shm_last_matmul_kernel_ms(void);
/* duration of the tile pass of the last batched matmul's fused rescale NTT, ms (HIP events) */
double shm_last_matmul_rescale_tile_ms(void);
/* variant the last matmul actually ran: SHM_MATMUL_INT8 falls back to SHM_MATMUL_U64
 * when a prime is outside (2^45, 2^62), a limb/weight digit pair has no compiled
 * contraction (6 or 7 limb digits with 4..6 weight digits, 8 with 4), or a row chunk's
 * sum could reach 2^106 (256 rows x 2^(8 digits - 1) x the largest |round(W scale)|). Any row_W is accepted: rows beyond 256 are
 * contracted in chunks (ct_pt_matrix_mul_wo_pre_large, Ct_pt_matrix_mul.cuh:584). */
int shm_last_matmul_variant(void);

/* ---- bootstrapping (slot-to-coefficient first) ------------------------------------
 * Replaces Bootstrapper (src/include/source/bootstrapping/Bootstrapper.cuh:63) and
 * the ordinary path of BatchBootstrapper (BatchBootstrap.cuh:470-820). Full slots
 * (n = N/2). Transform layers follow sfl_full_3 / coefftoslot_full_3
 * (Bootstrapper.cu:2737-2777, 3048-3067); EvalMod is a Chebyshev cosine of degree
 * `deg` on [-K, K] followed by `double_angles` double-angle steps. */
typedef struct shm_bootstrapper shm_bootstrapper;

typedef struct shm_boot_params {
    uint32_t s2c_input_primes; /* must be 4 (BatchBootstrap.cuh:727-728) */
    double layer3_factor;      /* F of the third S2C layer (kLowS2cLayer3Factor = 1/16) */
    int32_t K;                 /* EvalMod range: |integer part| < K (boundary_K = 25) */
    int32_t deg;               /* Chebyshev degree, 16..63 */
    int32_t double_angles;     /* r (num_double_formula = 2) */
    double input_bound;        /* bound on |Re z|, |Im z| of the slots (1.0) */
} shm_boot_params;

/* Bootstrapper ctor + generate_LT_coefficient_3 + prepare_mod_polynomial
 * (Bootstrapper.cu:3-20, 2135-2147; coefficients 634-727, 1356-2119); host only */
int shm_bootstrapper_create(shm_context *ctx, const shm_boot_params *params, shm_bootstrapper **out);
/* generate_LT_coefficient_3 + prepare_mod_polynomial alone (Bootstrapper.cu:2135-2147):
 * the host plan for N = 2^log_poly_degree and size_q data primes, without a context;
 * it can be inspected (steps, layers, diagonals, coefficients) but not run */
int shm_bootstrapper_create_plan(uint32_t log_poly_degree, uint32_t size_q, const shm_boot_params *params,
                                 shm_bootstrapper **out);
int shm_bootstrapper_destroy(shm_bootstrapper *b);
/* addBootKeys_3 (Bootstrapper.cu:456-479): rotation steps the transforms use, step 0 =
 * complex conjugation. Pass to shm_galois_keys_generate. `count` receives the total. */
int shm_bootstrapper_rotation_steps(const shm_bootstrapper *b, int *steps, size_t cap, size_t *count);
/* Layer schedule: transform 0 = slot-to-coefficient, 1 = coefficient-to-slot; layer
 * 0..2 in application order. info = {base step, k0, babies g, giants, chain_index}. */
int shm_bootstrapper_layer_info(const shm_bootstrapper *b, int transform, int layer, int64_t info[5]);
/* Pre-rotated diagonal (giant i, baby j) of a layer, n = N/2 complex values
 * (fftcoeff1/2/3 after rotation(); Bootstrapper.cu:2196-2197). Returns zeros for an
 * absent diagonal. */
int shm_bootstrapper_layer_diagonal(const shm_bootstrapper *b, int transform, int layer, int giant, int baby,
                                    double *re, double *im, size_t n);
/* EvalMod Chebyshev coefficients (deg + 1 values) */
int shm_bootstrapper_evalmod_coeffs(const shm_bootstrapper *b, double *out, size_t cap, size_t *count);
/* Encode every diagonal on the device at the scale of the prime its rescale drops
 * (BatchBootstrapper::prepare, BatchBootstrap.cuh:492). Synchronises the stream. */
int shm_bootstrapper_prepare(shm_bootstrapper *b, shm_stream_t stream);

/* Bootstrapper::modraise_inplace (Bootstrapper.cu:3287-3356): a 1-prime ciphertext
 * lifted to every data prime (centred: residues above q0/2 stand for x - q0). The
 * scale is left unchanged; callers set it to q0. */
int shm_modraise_inplace(shm_context *ctx, shm_ciphertext *ct, shm_stream_t stream);
/* Multiply by X^(N/2) (every slot times i), or by X^(3N/2) (times -i) when negate != 0:
 * the exact form of the encode(i) multiply_plain in coefftoslot_full_3 /
 * ordinary_evalmod_combine_batch (Bootstrapper.cu:3053-3062; BatchBootstrap.cuh:216-224) */
int shm_multiply_by_i_inplace(shm_context *ctx, shm_ciphertext *ct, int negate, shm_stream_t stream);
/* sfl_full_3 (Bootstrapper.cu:2737-2777) at the 4-prime level: 3 levels consumed */
int shm_bootstrap_slot_to_coeff(shm_bootstrapper *b, shm_ciphertext *ct, const shm_galois_keys *gk,
                                shm_stream_t stream);
/* coefftoslot_full_3 (Bootstrapper.cu:3048-3067): ct (top level, after modraise) ->
 * real and imaginary parts of its coefficient vector (bit-reversed), 3 levels */
int shm_bootstrap_coeff_to_slot(shm_bootstrapper *b, shm_ciphertext *ct, shm_ciphertext *re, shm_ciphertext *im,
                                const shm_galois_keys *gk, shm_stream_t stream);
/* ModularReducer::modular_reduction (ModularReducer.cu:64-84): slots x with
 * |x| < K -> sin(2 pi x), same scale; 1 + ceil(log2(deg+1)) + 1 + double_angles levels */
int shm_bootstrap_eval_mod(shm_bootstrapper *b, shm_ciphertext *ct, const shm_kswitch_key *relin,
                           shm_stream_t stream);
/* ordinary_evalmod_combine_batch (BatchBootstrap.cuh:192-260) runs EvalMod over a
 * chunk of ciphertexts: the same modular reduction on `count` (1..32) ciphertexts
 * at one level and scale, in lockstep, with one key-switch and rescale pass per
 * step for the group; each result equals shm_bootstrap_eval_mod's bit for bit. */
int shm_bootstrap_eval_mod_batch(shm_bootstrapper *b, shm_ciphertext *const *cts, size_t count,
                                 const shm_kswitch_key *relin, shm_stream_t stream);
/* ---- bootstrapping server (continuous batching) --------------------------------------
 * No counterpart in the reference, whose BatchBootstrapper (BatchBootstrap.cuh:470-820)
 * takes one fixed batch per call. Worker threads hold the keys resident and bootstrap
 * (S2C-first, as shm_bootstrap_s2c_first) the ciphertexts of all admitted requests in
 * groups of up to max_group (1..8) ciphertexts of one scale, slots of different requests
 * side by side; requests submitted while a group runs join the next group. Each output
 * equals, limb for limb, the bootstrap of that input alone.
 * Admission is by memory: a request enters while the estimated device footprint of the
 * admitted requests stays within the cap, the smaller of mem_budget_bytes (0: none) and
 * the device's free memory (hipMemGetInfo) less a reserve (one larger request enters
 * alone). With several workers (shm_server_create_ex, 1..4) each has its own stream
 * (default priority) and groups run concurrently. Inputs are copied at submit and
 * outputs delivered on the default stream. out[] must not be touched until the wait
 * returns for that request. Destroy finishes all submitted work. */
typedef struct shm_server shm_server;
typedef struct {
    uint64_t submitted, completed, failed, ciphertexts, groups;
    uint64_t in_flight_bytes;
    double p50_ms, p99_ms; /* request latency, submit to completion */
    double busy_s;         /* wall time inside grouped bootstraps, summed over workers */
    uint64_t packed_requests, shared_cts; /* packed requests done, shared ciphertexts bootstrapped */
    uint64_t free_bytes, admit_cap_bytes; /* device free memory and admission cap at the last admission */
    uint32_t workers;
} shm_server_stats;
int shm_server_create(shm_bootstrapper *b, const shm_galois_keys *gk, const shm_kswitch_key *relin,
                      uint64_t mem_budget_bytes, uint32_t max_group, shm_server **out);
int shm_server_create_ex(shm_bootstrapper *b, const shm_galois_keys *gk, const shm_kswitch_key *relin,
                         uint64_t mem_budget_bytes, uint32_t max_group, uint32_t workers, shm_server **out);
/* Packed request (the shared slot packing of uneven token counts): in[i] (count column
 * ciphertexts of one input in the batch_input layout, Batch_encode_encrypt.cuh:12-43)
 * holds its data in slots [0, slots) and zeros elsewhere. Requests of the same count and
 * scale are rotated to power-of-two-aligned slot blocks (the next power of two at or
 * above slots) and summed into shared ciphertexts, bootstrapped once for all of them;
 * out[i] receives the bootstrapped shared ciphertext i, with this request's slots at
 * [offset, offset + slots) (shm_server_wait_packed) and the other requests' data in the
 * other blocks (all requests share the secret key). The Galois keys must hold the
 * steps +-2^k for every power of two from the block up to N/4. */
int shm_server_submit_packed(shm_server *srv, const shm_ciphertext *const *in, size_t count, uint32_t slots,
                             shm_ciphertext *const *out, uint64_t *request_id);
int shm_server_wait_packed(shm_server *srv, uint64_t request_id, double *latency_ms, uint32_t *slot_offset);
int shm_server_destroy(shm_server *srv);
int shm_server_submit(shm_server *srv, const shm_ciphertext *const *in, size_t count, shm_ciphertext *const *out,
                      uint64_t *request_id);
int shm_server_wait(shm_server *srv, uint64_t request_id, double *latency_ms);
int shm_server_get_stats(shm_server *srv, shm_server_stats *stats);

/* Polynomial::homomorphic_poly_evaluation (bootstrapping/common/Polynomial.cu:253-476)
 * in Chebyshev form, as the non-linear layers use it (gelu_cheb,
 * non_linear_func/gelu_other.cuh:257-330: sum_k d_k T_k(u), u affine in x):
 * each of `count` (1..32) ciphertexts at one level and scale becomes
 * sum_{k < n_coeffs} coeffs[k] T_k((2x - lo - hi) / (hi - lo)) in place, at the input
 * scale. Degree n_coeffs - 1 in 1..63; baby steps T_1..T_8, giant steps T_8, T_16
 * and T_32; depth 1 (affine map) + ceil(log2(n_coeffs)) + 1 levels at most. All
 * ciphertexts go through each product together (one key switch and rescale pass). */
int shm_evaluate_chebyshev(shm_context *ctx, shm_ciphertext *const *cts, size_t count, const double *coeffs,
                           size_t n_coeffs, double lo, double hi, const shm_kswitch_key *relin, shm_stream_t stream);
/* Ct_ct_matrix_mul.cuh:10-181 ct_ct_matrix_mul_colpacking: out[i] (i < row_X) =
 * rescale(relinearize(sum_{j < col_X} enc_X[j] * rotate(enc_W[j], i * num_batch))), the
 * diagonal-packed product of two column-packed encrypted matrices. col_X must equal
 * col_W and row_X row_W (the reference's check); all operands share one level, the X
 * operands one scale and the W operands one scale; outputs must not alias inputs. */
int shm_ct_ct_matrix_mul_colpacking(shm_context *ctx, const shm_ciphertext *const *enc_X,
                                    const shm_ciphertext *const *enc_W, int col_X, int row_X, int col_W, int row_W,
                                    int num_batch, const shm_galois_keys *gk, const shm_kswitch_key *relin,
                                    shm_ciphertext *const *out, shm_stream_t stream);
/* Ct_ct_matrix_mul.cuh:182-609 ct_ct_matrix_mul_diagpacking (baby-step giant-step,
 * g = ceil(sqrt(col_X)) babies): out[i] (i < col_W) = sum_j rotate(giant_j, j g num_batch)
 * with giant_j = rescale(relinearize(sum_k rotate(enc_W[i], k num_batch) *
 * rotate(enc_X[j g + k], (col_X - j g) num_batch))), labelled with enc_X's scale as in
 * the reference. Reads enc_X[0 .. col_X); row_X must be at least col_X (the reference
 * indexes its pre-rotated X by col_X); row_W is not used, as in the reference. */
int shm_ct_ct_matrix_mul_diagpacking(shm_context *ctx, const shm_ciphertext *const *enc_X,
                                     const shm_ciphertext *const *enc_W, int col_X, int row_X, int col_W, int row_W,
                                     int num_batch, const shm_galois_keys *gk, const shm_kswitch_key *relin,
                                     shm_ciphertext *const *out, shm_stream_t stream);
/* BatchBootstrapper::bootstrap_ordinary_s2c_first (BatchBootstrap.cuh:674-709) over
 * `count` ciphertexts with exactly 4 primes; out[i] receives the refreshed
 * ciphertext at scale out_scale (0 = the input scale). Inputs are not modified.
 * Synchronises the stream. Outputs keep their buffers across calls (the result is
 * copied into them). */
int shm_bootstrap_s2c_first(shm_bootstrapper *b, const shm_ciphertext *const *in, size_t count,
                            shm_ciphertext *const *out, const shm_galois_keys *gk, const shm_kswitch_key *relin,
                            double out_scale, shm_stream_t stream);

/* ---- shared-a batches (SharedACiphertextBatch, SharedAKeyGen, FmtSwitch, Sh-Rot) ------
 * A batch holds n members (b_i, a) that share the a-part; member i decrypts under the
 * auxiliary secret s'_i. Batches of 1..16 members. */
typedef struct shm_shared_a_keys shm_shared_a_keys;
typedef struct shm_shared_a_batch shm_shared_a_batch;

/* SharedAKeyGen(context, sk, n) + gen_format_key(chain_index) (FmtSwitch.cu:154-270):
 * auxiliary secrets (Hamming weight hw) from splitmix64(seed, i), the format key at
 * chain_index (the bootstrap input level) and the backward keys s'_i -> sk. Synchronises. */
int shm_shared_a_keys_create(shm_context *ctx, const shm_secret_key *sk, uint32_t n, uint32_t hw, uint64_t seed,
                             uint32_t chain_index, shm_shared_a_keys **out);
int shm_shared_a_keys_destroy(shm_shared_a_keys *keys);
/* gen_shared_galois_key / acquire_shared_galois_key (ShRotate.cu:54-195) for each step
 * (0 = conjugation); keys stay resident (the reference's amortised mode). Synchronises. */
int shm_shared_a_keys_add_galois(shm_shared_a_keys *keys, const int *steps, size_t n_steps, uint64_t seed);
/* auxiliary secret i as N signed coefficients */
int shm_shared_a_keys_aux_coeffs(const shm_shared_a_keys *keys, uint32_t i, int8_t *out, size_t n);
/* format key words: a_fmt n x (Ql+P) x N, B n x n x (Ql+P) x N (index j*n + i), local limb order */
int shm_shared_a_keys_export_format(const shm_shared_a_keys *keys, uint64_t *a_fmt, uint64_t *B, size_t n_words_B);
/* shared Galois key of element elt: A dnum x QP x N, B dnum x n x QP x N */
int shm_shared_a_keys_export_galois(const shm_shared_a_keys *keys, uint32_t elt, uint64_t *A, uint64_t *B,
                                    size_t n_words_B);

int shm_shared_a_batch_create(shm_context *ctx, uint32_t n, uint32_t chain_index, shm_shared_a_batch **out);
int shm_shared_a_batch_destroy(shm_shared_a_batch *b);
int shm_shared_a_batch_info(const shm_shared_a_batch *b, uint32_t *n, uint32_t *chain_index,
                            uint32_t *coeff_mod_size, double *scale);
int shm_shared_a_batch_set_scale(shm_shared_a_batch *b, double scale);
/* device words [n+1][Ql][N]: b_0..b_{n-1}, then a (the public a_ / b_ buffers
 * of SharedACiphertextBatch, SharedACiphertext.cuh:29-30, in one allocation) */
uint64_t *shm_shared_a_batch_data(shm_shared_a_batch *b);
/* words [n+1][Ql][N]: b_0..b_{n-1}, then a */
int shm_shared_a_batch_import(shm_shared_a_batch *b, const uint64_t *host, size_t n_words, shm_stream_t stream);
int shm_shared_a_batch_export(const shm_shared_a_batch *b, uint64_t *host, size_t n_words, shm_stream_t stream);
/* SharedACiphertextBatch::assemble (SharedACiphertext.cuh:74-99): member i as (b_i, a) */
int shm_shared_a_batch_member(const shm_shared_a_batch *b, uint32_t i, shm_ciphertext *out, shm_stream_t stream);

/* fmt_switch_forward (FmtSwitch.cu:276-394): n ciphertexts at the key's level -> batch */
int shm_fmt_switch_forward(shm_shared_a_keys *keys, const shm_ciphertext *const *cts, size_t n,
                           shm_shared_a_batch *out, shm_stream_t stream);
/* fmt_switch_backward (FmtSwitch.cu:396-429): batch -> n ciphertexts under sk */
int shm_fmt_switch_backward(shm_shared_a_keys *keys, const shm_shared_a_batch *batch, shm_ciphertext *const *out,
                            size_t n, shm_stream_t stream);
/* shared_a_rotate (ShRotate.cu:201-257) by a slot step (0 = conjugation) */
int shm_shared_a_rotate_inplace(shm_shared_a_keys *keys, shm_shared_a_batch *b, int step, shm_stream_t stream);
/* SharedABatchOps.cu:16-166: add, multiply_plain, rescale_to_next, modraise */
int shm_shared_a_add_inplace(shm_shared_a_batch *a, const shm_shared_a_batch *b, shm_stream_t stream);
int shm_shared_a_multiply_plain_inplace(shm_shared_a_batch *b, const shm_plaintext *pt, shm_stream_t stream);
int shm_shared_a_rescale_to_next_inplace(shm_shared_a_batch *b, shm_stream_t stream);
int shm_shared_a_modraise_inplace(shm_shared_a_batch *b, shm_stream_t stream);
/* Bootstrapper::shared_a_sfl_full_3 (SharedACoeffToSlot.cu:311-358) on a batch at the
 * 4-prime level (3 levels consumed) */
int shm_bootstrap_shared_slot_to_coeff(shm_bootstrapper *b, shm_shared_a_keys *keys, shm_shared_a_batch *batch,
                                       shm_stream_t stream);
/* Bootstrapper::shared_a_coefftoslot_full_3 (SharedACoeffToSlot.cu:269-301): batch at the
 * top level (after shm_shared_a_modraise_inplace) -> real / imaginary part batches */
int shm_bootstrap_shared_coeff_to_slot(shm_bootstrapper *b, shm_shared_a_keys *keys, shm_shared_a_batch *batch,
                                       shm_shared_a_batch *re, shm_shared_a_batch *im, shm_stream_t stream);
/* BatchBootstrapper::bootstrap_shared_a_s2c_first (BatchBootstrap.cuh:719-820): exactly n
 * ciphertexts with 4 primes (n = the keys' batch size); out_scale 0 = the input scale.
 * Synchronises the stream. */
int shm_bootstrap_shared_a_s2c_first(shm_bootstrapper *b, shm_shared_a_keys *keys, const shm_ciphertext *const *in,
                                     size_t n, shm_ciphertext *const *out, const shm_kswitch_key *relin,
                                     double out_scale, shm_stream_t stream);

/* BatchBootstrapper::bootstrap_shared_a_s2c_first over count = k n ciphertexts, k chunks of
 * the keys' batch size n in order (the driver loop of test_shared_a.cuh:1336-1392).
 * Synchronises the stream once, after the last chunk. */
int shm_bootstrap_shared_a_s2c_first_chunks(shm_bootstrapper *b, shm_shared_a_keys *keys,
                                            const shm_ciphertext *const *in, size_t count,
                                            shm_ciphertext *const *out, const shm_kswitch_key *relin,
                                            double out_scale, shm_stream_t stream);

/* ---- graph-captured bootstrapping (no counterpart in the reference, whose driver loop
 * launches every kernel of every chunk from the host) ------------------------------------
 * One chunk is recorded into a HIP graph on `stream` (a stream from shm_stream_create): a
 * shared-a batch of n = the keys' batch size, or an ordinary group of `group` (1..8)
 * ciphertexts. Creation runs the chunk once eagerly on zero inputs, then records it.
 * shm_boot_graph_run bootstraps count (a multiple of the chunk) ciphertexts with the
 * recorded level and scale: per chunk it copies the inputs into the graph's input
 * buffers, replays the graph and copies the outputs out, all on the recording stream,
 * ordered after the work queued on the default stream and before later default-stream
 * work; it only enqueues. The bootstrapper and keys must outlive the graph.
 * Currently switched off: creation returns SHM_ERR_LOGIC (the first replay of a
 * recorded chunk on an MI355X faulted; the cause is not established). */
typedef struct shm_boot_graph shm_boot_graph;
int shm_boot_graph_create_shared_a(shm_bootstrapper *b, shm_shared_a_keys *keys, const shm_kswitch_key *relin,
                                   double in_scale, double out_scale, shm_stream_t stream, shm_boot_graph **out);
int shm_boot_graph_create_ordinary(shm_bootstrapper *b, const shm_galois_keys *gk, const shm_kswitch_key *relin,
                                   uint32_t group, double in_scale, double out_scale, shm_stream_t stream,
                                   shm_boot_graph **out);
int shm_boot_graph_run(shm_boot_graph *g, const shm_ciphertext *const *in, size_t count, shm_ciphertext *const *out);
/* chunk size and the number of recorded graph nodes */
int shm_boot_graph_info(const shm_boot_graph *g, uint32_t *chunk, uint64_t *nodes);
int shm_boot_graph_destroy(shm_boot_graph *g);

/* ---- measurement hooks (no counterpart in the reference, which times no kernel) -------
 * Probe of the fused last NTT pass (the tile pass that writes the ModDown combine
 * (cx - NTT(delta)) P^-1 and the rescale (acc - NTT(last)) q_last^-1): between begin and
 * end every eager launch is bracketed by timing events on its stream; end waits for
 * the device and returns the launch count, the summed kernel milliseconds and the
 * algorithmic bytes of those launches (per output limb: the result written and the
 * combined operand read, plus the output read where it accumulates). */
int shm_probe_fused_ntt_begin(void);
int shm_probe_fused_ntt_end(uint64_t *launches, double *ms, double *bytes);
/* The plain forward NTT tile passes (the ModUp digits' and other forward transforms' last
 * pass) timed in the same window, with 16 N algorithmic bytes per limb (each limb read and
 * written once); valid after shm_probe_fused_ntt_end. */
int shm_probe_plain_ntt_last(uint64_t *launches, double *ms, double *bytes);
/* Key-switching work since the last reset, per level Ql (entry Ql, cap entries): polynomials
 * ModUp'd, inner-product output polynomials and polynomials ModDown'd. */
int shm_keyswitch_counters(uint64_t *modup, uint64_t *inner, uint64_t *moddown, size_t cap);
int shm_keyswitch_counters_reset(void);
/* Launch-shape knobs (no counterpart in the reference; results never depend on them):
 * "ks_batch_slots_log2" and "ks_multi_slots_log2" set log2 of the member / output slots
 * per block of the batched and multi-output key inner products (0..4), or -1 (the
 * default) for the autotuned choice; "matmul_overlap" (0, the default, or 1) runs the INT8
 * ct-pt matmul's contraction and rescale one after the other, or overlaps the contraction
 * (row_W <= 256, in limb windows) with the rescale on a second stream of the context;
 * "ntt_tile_threads" (512, the default, or 256) runs the NTT tile passes in 512-thread
 * blocks of radix-8 rounds or 256-thread blocks of radix-16 rounds (the same words).
 * Process-wide. */
int shm_set_tuning(const char *name, int32_t value);
int shm_get_tuning(const char *name, int32_t *value);
/* Autotuner (no counterpart in the reference): the first call of each inner-product
 * shape (kernel, N, Ql, members or outputs, digits, permuted, ...) that writes its
 * output outright times every candidate slot count on its own operands (it waits for
 * its stream once) and caches the fastest; calls recorded into a graph use the cached
 * choice or one slot. On by default. The report lists the cached shapes, the choice and
 * the timings, one line each (needed = bytes including the terminating zero). */
int shm_autotune_enable(int32_t on);
int shm_autotune_clear(void);
int shm_autotune_report(char *buf, size_t cap, size_t *needed);

#ifdef __cplusplus
}
#endif

#endif /* SHMOAI_H */
