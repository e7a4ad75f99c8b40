// Elementwise RNS kernels, rescale, Galois permutation, samplers and key/
// encryption combiners. Grid-stride loops over [poly][limb][N], one
// coefficient per thread per iteration.
#include "engine.h"

namespace shm {

namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(size_t work) {
    size_t g = (work + kBlock - 1) / kBlock;
    if (g > (1u << 20)) g = 1u << 20;
    return (unsigned)(g ? g : 1);
}

// Row grid for [rows][N] work: blockIdx.y walks the rows (one limb of one
// polynomial each), blockIdx.x and the thread walk the N words of a row, so
// the limb of a row is one uniform value per block rather than a 64-bit
// division per word. Threads take 4 words of a row when the launch is large.
inline dim3 rows_grid(size_t rows, size_t N) {
    size_t bx = (N + kBlock - 1) / kBlock;
    if (rows * bx >= 8192) bx = (bx + 3) / 4;
    const size_t by = rows == 0 ? 1 : (rows < 65535 ? rows : 65535);
    return dim3((unsigned)bx, (unsigned)by);
}

#define SHM_ROWS_LOOP(row, rows) for (size_t row = blockIdx.y; row < (rows); row += gridDim.y)
#define SHM_COLS_LOOP(i, N) for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (N); i += (size_t)gridDim.x * blockDim.x)

__global__ void k_add(uint64_t *a, const uint64_t *b, const DevMod *mods, size_t Ql, size_t N, size_t rows) {
    SHM_ROWS_LOOP(row, rows) {
        const uint64_t q = mods[row % Ql].q;
        uint64_t *ar = a + row * N;
        const uint64_t *br = b + row * N;
        SHM_COLS_LOOP(i, N) ar[i] = add_mod(ar[i], br[i], q);
    }
}

__global__ void k_sub(uint64_t *a, const uint64_t *b, const DevMod *mods, size_t Ql, size_t N, size_t rows) {
    SHM_ROWS_LOOP(row, rows) {
        const uint64_t q = mods[row % Ql].q;
        uint64_t *ar = a + row * N;
        const uint64_t *br = b + row * N;
        SHM_COLS_LOOP(i, N) ar[i] = sub_mod(ar[i], br[i], q);
    }
}

__global__ void k_negate(uint64_t *a, const DevMod *mods, size_t Ql, size_t N, size_t rows) {
    SHM_ROWS_LOOP(row, rows) {
        const uint64_t q = mods[row % Ql].q;
        uint64_t *ar = a + row * N;
        SHM_COLS_LOOP(i, N) ar[i] = neg_mod(ar[i], q);
    }
}

// polymath.cu:161 multiply_rns_poly: ct (.) pt per limb, pt broadcast over polys.
__global__ void k_mul_plain(uint64_t *ct, const uint64_t *pt, const DevMod *mods, size_t Ql, size_t N, size_t rows) {
    SHM_ROWS_LOOP(row, rows) {
        const size_t limb = row % Ql;
        const DevMod m = mods[limb];
        uint64_t *cr = ct + row * N;
        const uint64_t *pr = pt + limb * N;
        SHM_COLS_LOOP(i, N) cr[i] = mul_mod(cr[i], pr[i], m);
    }
}

// polymath.cu:468-503 tensor_prod_2x2_rns_poly (Karatsuba middle term).
// a and b hold a_limbs and b_limbs limbs per polynomial (>= Ql): an operand at a
// higher level is used through its first Ql limbs, which is what dropping its
// top limbs (mod_switch_to) would leave. One row per limb. With `accumulate`
// the three products are added onto out (a sum of tensor products, as
// multiply + add_inplace of 3-polynomial ciphertexts).
__global__ void k_tensor(const uint64_t *a, const uint64_t *b, uint64_t *out, const DevMod *mods, size_t Ql,
                         size_t N, size_t a_limbs, size_t b_limbs, int accumulate, int p52) {
    const size_t S = Ql * N, Sa = a_limbs * N, Sb = b_limbs * N;
    SHM_ROWS_LOOP(row, Ql) {
        const DevMod m = mods[row];
        SHM_COLS_LOOP(i, N) {
            const size_t t = row * N + i;
            const uint64_t a0 = a[t], a1 = a[t + Sa], b0 = b[t], b1 = b[t + Sb];
            // 52-bit primes: split products (operands below 2q < 2^53), same residues
            auto mm = [&](uint64_t x, uint64_t y) { return p52 ? mul_mod52(x, y, m) : mul_mod(x, y, m); };
            const uint64_t d0 = mm(a0, b0), d2 = mm(a1, b1);
            uint64_t d1 = mm(a0 + a1, b0 + b1);
            d1 = d1 + 2 * m.q - d0 - d2;
            d1 = csub(d1, m.q);
            d1 = csub(d1, m.q);
            if (accumulate) {
                out[t] = add_mod(out[t], d0, m.q);
                out[t + S] = add_mod(out[t + S], d1, m.q);
                out[t + 2 * S] = add_mod(out[t + 2 * S], d2, m.q);
            } else {
                __builtin_nontemporal_store(d0, &out[t]);
                __builtin_nontemporal_store(d1, &out[t + S]);
                __builtin_nontemporal_store(d2, &out[t + 2 * S]);
            }
        }
    }
}

// One step of the Chebyshev basis recursion T_k = 2 T_a T_b - T_(a-b), before the
// relinearisation: (d0, d1, d2) = 2 (a (x) b) - cval (t0, t1, 0) with cval an integer
// (reduced per prime here), so that after the rescale the T_(a-b) term carries the
// product's scale; t == nullptr leaves the doubled product; b == a squares. One row per
// limb; every operand enters through its first Ql limbs.
__global__ void k_tensor_cheb(const uint64_t *a, const uint64_t *b, const uint64_t *t, uint64_t *out,
                              const DevMod *mods, size_t Ql, size_t N, size_t a_limbs, size_t b_limbs, size_t t_limbs,
                              uint64_t cval, int p52) {
    const size_t S = Ql * N, Sa = a_limbs * N, Sb = b_limbs * N, St = t_limbs * N;
    SHM_ROWS_LOOP(row, Ql) {
        const DevMod m = mods[row];
        const uint64_t q = m.q;
        const uint64_t cm = reduce64(cval, m), cms = shoup_companion(cm, m);
        SHM_COLS_LOOP(i, N) {
            const size_t x = row * N + i;
            const uint64_t a0 = a[x], a1 = a[x + Sa];
            // 52-bit primes: split products (operands below 2q < 2^53), same residues
            auto mm = [&](uint64_t u, uint64_t w) { return p52 ? mul_mod52(u, w, m) : mul_mod(u, w, m); };
            uint64_t d0, d1, d2;
            if (a == b) {
                d0 = mm(a0, a0);
                d2 = mm(a1, a1);
                if (p52) {
                    d1 = mul_mod52(a0, a1 << 1, m);
                } else {
                    const uint64_t lo = a0 * a1, hi = __umul64hi(a0, a1);
                    d1 = barrett128((hi << 1) | (lo >> 63), lo << 1, m);
                }
            } else {
                const uint64_t b0 = b[x], b1 = b[x + Sb];
                d0 = mm(a0, b0);
                d2 = mm(a1, b1);
                d1 = mm(a0 + a1, b0 + b1);
                d1 = csub(csub(d1 + 2 * q - d0 - d2, q), q);
            }
            d0 = add_mod(d0, d0, q);
            d1 = add_mod(d1, d1, q);
            d2 = add_mod(d2, d2, q);
            if (t) {
                d0 = sub_mod(d0, mul_shoup(t[x], cm, cms, q), q);
                d1 = sub_mod(d1, mul_shoup(t[x + St], cm, cms, q), q);
            }
            __builtin_nontemporal_store(d0, &out[x]);
            __builtin_nontemporal_store(d1, &out[x + S]);
            __builtin_nontemporal_store(d2, &out[x + 2 * S]);
        }
    }
}

// polymath.cu:505-550 tensor_square_2x2_rns_poly. One row per limb.
__global__ void k_square(const uint64_t *a, uint64_t *out, const DevMod *mods, size_t Ql, size_t N,
                         size_t a_limbs, int p52) {
    const size_t S = Ql * N, Sa = a_limbs * N;
    SHM_ROWS_LOOP(row, Ql) {
        const DevMod m = mods[row];
        SHM_COLS_LOOP(i, N) {
            const size_t t = row * N + i;
            const uint64_t c0 = a[t], c1 = a[t + Sa];
            uint64_t d0, d1, d2;
            if (p52) {
                // split products (2 c1 < 2^53), same residues
                d0 = mul_mod52(c0, c0, m);
                d2 = mul_mod52(c1, c1, m);
                d1 = mul_mod52(c0, c1 << 1, m);
            } else {
                d0 = mul_mod(c0, c0, m);
                d2 = mul_mod(c1, c1, m);
                const uint64_t lo = c0 * c1, hi = __umul64hi(c0, c1);
                d1 = barrett128((hi << 1) | (lo >> 63), lo << 1, m);
            }
            out[t] = d0;
            out[t + S] = d1;
            out[t + 2 * S] = d2;
        }
    }
}

// limb `limb` of every polynomial [npolys][Ql][N] -> [npolys][N]
__global__ void k_gather_limb(const uint64_t *in, uint64_t *out, size_t Ql, size_t limb, size_t N, size_t npolys) {
    const size_t total = npolys * N;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t p = t >> __builtin_ctzll(N), j = t & (N - 1);
        out[t] = in[(p * Ql + limb) * N + j];
    }
}

// out[p][i] = in.p[p][i] for i < words: polynomials of separate buffers into one
// (one row per polynomial).
__global__ void k_gather_polys(PolyPack in, uint64_t *__restrict__ out, size_t words, size_t npolys) {
    SHM_ROWS_LOOP(p, npolys) {
        const uint64_t *src = in.p[p];
        uint64_t *dst = out + p * words;
        SHM_COLS_LOOP(i, words) __builtin_nontemporal_store(src[i], &dst[i]);
    }
}

// galois.cu:8-20 apply_galois_ntt_permutation (one row per limb).
__global__ void k_permute(const uint64_t *in, uint64_t *out, const uint32_t *table, size_t N, size_t nlimbs) {
    SHM_ROWS_LOOP(limb, nlimbs) {
        const uint64_t *src = in + limb * N;
        uint64_t *dst = out + limb * N;
        SHM_COLS_LOOP(i, N) __builtin_nontemporal_store(src[table[i]], &dst[i]);
    }
}

// The first out_limbs limbs of each polynomial (one row per output limb).
__global__ void k_copy_limbs(const uint64_t *in, size_t in_limbs, uint64_t *out, size_t out_limbs, size_t N,
                             size_t npolys) {
    SHM_ROWS_LOOP(row, npolys * out_limbs) {
        const size_t p = row / out_limbs, l = row % out_limbs;
        const uint64_t *src = in + (p * in_limbs + l) * N;
        uint64_t *dst = out + row * N;
        SHM_COLS_LOOP(i, N) dst[i] = src[i];
    }
}

// a (+)= / (*)= the constant [round(value * scale)]_q of every limb, the residue
// being what encode_scalar writes into each slot of the constant plaintext (so
// the result equals add_plain / multiply_plain with that plaintext).
// One row per limb; the residue (and for the product its Shoup companion) is
// formed once per row.
__global__ void k_add_scalar(uint64_t *a, const DevMod *mods, double v, size_t N, size_t Ql) {
    SHM_ROWS_LOOP(row, Ql) {
        const DevMod m = mods[row];
        uint64_t r = residue_of_rounded(v, m);
        r = r >= m.q ? r - m.q : r;
        uint64_t *ar = a + row * N;
        SHM_COLS_LOOP(i, N) ar[i] = add_mod(ar[i], r, m.q);
    }
}

// The same over separate polynomials: row r is limb r % Ql of polynomial a.p[r / Ql];
// op 0: a += [v]_q, op 1: a *= [v]_q, op 2: a += b (b.p[r / Ql]).
__global__ void k_pack_op(PolyPack a, PolyPack b, const DevMod *mods, double v, size_t Ql, size_t N, size_t rows,
                          int op) {
    SHM_ROWS_LOOP(row, rows) {
        const size_t p = row / Ql, k = row % Ql;
        const DevMod m = mods[k];
        uint64_t *ar = a.p[p] + k * N;
        if (op == 2) {
            const uint64_t *br = b.p[p] + k * N;
            SHM_COLS_LOOP(i, N) ar[i] = add_mod(ar[i], br[i], m.q);
            continue;
        }
        uint64_t r = residue_of_rounded(v, m);
        r = r >= m.q ? r - m.q : r;
        if (op == 0) {
            SHM_COLS_LOOP(i, N) ar[i] = add_mod(ar[i], r, m.q);
        } else {
            const uint64_t rs = shoup_companion(r, m);
            SHM_COLS_LOOP(i, N) ar[i] = mul_shoup(ar[i], r, rs, m.q);
        }
    }
}

__global__ void k_mul_scalar(uint64_t *a, const DevMod *mods, double v, size_t Ql, size_t N, size_t rows) {
    SHM_ROWS_LOOP(row, rows) {
        const DevMod m = mods[row % Ql];
        uint64_t r = residue_of_rounded(v, m);
        r = r >= m.q ? r - m.q : r;
        const uint64_t rs = shoup_companion(r, m);
        uint64_t *ar = a + row * N;
        SHM_COLS_LOOP(i, N) ar[i] = mul_shoup(ar[i], r, rs, m.q);
    }
}

__global__ void k_fill_zero(uint64_t *out, size_t words) {
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < words; t += (size_t)gridDim.x * blockDim.x)
        out[t] = 0;
}

// ---- samplers (SplitMix64 counters; identical streams on host and device) --
__device__ __forceinline__ uint64_t splitmix64(uint64_t key, uint64_t ctr) {
    uint64_t z = key + (ctr + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__device__ __forceinline__ uint32_t limb_prime(size_t limb, size_t Ql, size_t nQ) {
    return (uint32_t)(limb < Ql ? limb : nQ + (limb - Ql));
}

// Uniform residue per (prime, coefficient): (draw0 * 2^64 + draw1) mod q.
__global__ void k_sample_uniform(uint64_t seed, uint64_t *out, const DevMod *mods, size_t Ql, size_t nQ, size_t N,
                                 size_t total) {
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t limb = t >> __builtin_ctzll(N), j = t & (N - 1);
        const uint32_t pi = limb_prime(limb, Ql, nQ);
        const uint64_t ctr = 2 * ((uint64_t)pi * N + j);
        out[t] = barrett128(splitmix64(seed, ctr), splitmix64(seed, ctr + 1), mods[pi]);
    }
}

// Discrete Gaussian sigma = 3.2 via its cumulative table (|e| <= 19): the
// 63 high bits of a draw select the magnitude, bit 0 the sign.
__constant__ uint64_t kCDT[19] = {
    0x0ff52b40a5917f1dULL, 0x2e5a25d4bf0e400eULL, 0x489ae26955b04bd6ULL, 0x5d2bc20f621bf185ULL,
    0x6bc8694c3cc80ff4ULL, 0x7532d89ac6ba7dceULL, 0x7ab396cb74436798ULL, 0x7d9e4e916643eb07ULL,
    0x7f05495819eb2051ULL, 0x7fa1ce9c0039a957ULL, 0x7fdfb3f212e8c4e8ULL, 0x7ff5e6f9d2314fccULL,
    0x7ffd1f97bc4406a2ULL, 0x7fff40fa0088d11dULL, 0x7fffd2e835e1c57dULL, 0x7ffff6524386ff1eULL,
    0x7ffffe1db4769da5ULL, 0x7fffffac0a1dcb08ULL, 0x7ffffff428673853ULL};

__global__ void k_sample_gaussian(uint64_t seed, uint64_t *out, const DevMod *mods, size_t Ql, size_t nQ, size_t N,
                                  size_t total) {
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t limb = t >> __builtin_ctzll(N), j = t & (N - 1);
        const uint64_t u = splitmix64(seed, j), v = u >> 1;
        uint32_t mag = 0;
#pragma unroll
        for (int i = 0; i < 19; i++) mag += v >= kCDT[i] ? 1u : 0u;
        const uint64_t q = mods[limb_prime(limb, Ql, nQ)].q;
        out[t] = (u & 1) ? (mag ? q - mag : 0) : mag;
    }
}

__global__ void k_signed_to_rns(const int8_t *coeffs, uint64_t *out, const DevMod *mods, size_t Ql, size_t nQ,
                                size_t N, size_t total) {
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t limb = t >> __builtin_ctzll(N), j = t & (N - 1);
        const int64_t v = coeffs[j];
        const uint64_t q = mods[limb_prime(limb, Ql, nQ)].q;
        out[t] = v >= 0 ? (uint64_t)v : q - (uint64_t)(-v);
    }
}

// secretkey.cu:328-373: b = -(a s + e) (+ P * newkey on the digit's limbs), over QP.
__global__ void k_ks_key_digit(const uint64_t *sk, const uint64_t *newkey, uint64_t *b, const uint64_t *a,
                               const uint64_t *e, const DevMod *mods, const uint64_t *pmod, const uint64_t *pmod_s,
                               size_t dlo, size_t dhi, size_t N, size_t total) {
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t limb = t >> __builtin_ctzll(N);
        const DevMod m = mods[limb];
        uint64_t v = neg_mod(add_mod(mul_mod(a[t], sk[t], m), e[t], m.q), m.q);
        if (limb >= dlo && limb < dhi) v = add_mod(v, mul_shoup(newkey[t], pmod[limb], pmod_s[limb], m.q), m.q);
        b[t] = v;
    }
}

// secretkey.cu:260-326 + add pt: c0 = -(a s + e) + m.
__global__ void k_encrypt_combine(const uint64_t *sk, const uint64_t *a, const uint64_t *e, const uint64_t *pt,
                                  uint64_t *c0, const DevMod *mods, size_t N, size_t total) {
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const DevMod m = mods[t >> __builtin_ctzll(N)];
        const uint64_t v = neg_mod(add_mod(mul_mod(a[t], sk[t], m), e[t], m.q), m.q);
        c0[t] = add_mod(v, pt[t], m.q);
    }
}

// secretkey.cu:725-766: m = c0 + c1 s (+ c2 s^2).
__global__ void k_decrypt(const uint64_t *ct, size_t size, const uint64_t *sk, uint64_t *pt, const DevMod *mods,
                          size_t N, size_t S) {
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < S; t += (size_t)gridDim.x * blockDim.x) {
        const DevMod m = mods[t >> __builtin_ctzll(N)];
        const uint64_t s = sk[t];
        uint64_t v = add_mod(ct[t], mul_mod(ct[S + t], s, m), m.q);
        if (size > 2) v = add_mod(v, mul_mod(ct[2 * S + t], mul_mod(s, s, m), m), m.q);
        pt[t] = v;
    }
}

__global__ void k_mul_elementwise(const uint64_t *a, const uint64_t *b, uint64_t *out, const DevMod *mods, size_t Ql,
                                  size_t nQ, size_t N, size_t total) {
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const DevMod m = mods[limb_prime((t >> __builtin_ctzll(N)), Ql, nQ)];
        out[t] = mul_mod(a[t], b[t], m);
    }
}

// (splitmix64(seed, j) mod 3) - 1 over every limb (the role of sample_ternary_poly).
__global__ void k_sample_ternary_uniform(uint64_t seed, uint64_t *out, const DevMod *mods, size_t Ql, size_t nQ,
                                         size_t N, size_t total) {
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t limb = t >> __builtin_ctzll(N), j = t & (N - 1);
        const uint64_t r = splitmix64(seed, j) % 3;
        const uint64_t q = mods[limb_prime(limb, Ql, nQ)].q;
        out[t] = r == 0 ? q - 1 : r - 1;
    }
}

// secretkey.cu:15-85 (NTT branch): c'_p = u * pk_p + e over QP; npolys outputs.
__global__ void k_pk_mac(const uint64_t *u, const uint64_t *pk, const uint64_t *e, uint64_t *out, const DevMod *mods,
                         size_t N, size_t S, size_t total) {
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t w = t % S;
        const DevMod m = mods[w >> __builtin_ctzll(N)];
        out[t] = add_mod(mul_mod(u[w], pk[t], m), e[w], m.q);
    }
}

}  // namespace

void launch_sample_ternary_uniform(const Context &c, uint64_t seed, uint64_t *out, size_t Ql, bool withP,
                                   hipStream_t s) {
    const size_t limbs = Ql + (withP ? c.nP : 0), total = limbs * c.N;
    k_sample_ternary_uniform<<<grid_for(total), kBlock, 0, s>>>(seed, out, c.mods(), Ql, c.nQ, c.N, total);
    SHM_HIP(hipGetLastError());
}

void launch_neg_as_plus_e(const Context &c, const uint64_t *sk, const uint64_t *a, const uint64_t *e, uint64_t *b,
                          hipStream_t s) {
    const size_t total = c.nQP() * c.N;
    k_ks_key_digit<<<grid_for(total), kBlock, 0, s>>>(sk, nullptr, b, a, e, c.mods(), c.pmod.as<uint64_t>(),
                                                      c.pmod_s.as<uint64_t>(), 0, 0, c.N, total);
    SHM_HIP(hipGetLastError());
}

void launch_pk_mac(const Context &c, const uint64_t *u, const uint64_t *pk, const uint64_t *e, uint64_t *out,
                   size_t npolys, hipStream_t s) {
    const size_t S = c.nQP() * c.N, total = npolys * S;
    k_pk_mac<<<grid_for(total), kBlock, 0, s>>>(u, pk, e, out, c.mods(), c.N, S, total);
    SHM_HIP(hipGetLastError());
}

void launch_add(const Context &c, uint64_t *a, const uint64_t *b, size_t Ql, size_t npolys, hipStream_t s) {
    k_add<<<rows_grid(npolys * Ql, c.N), kBlock, 0, s>>>(a, b, c.mods(), Ql, c.N, npolys * Ql);
    SHM_HIP(hipGetLastError());
}

void launch_sub(const Context &c, uint64_t *a, const uint64_t *b, size_t Ql, size_t npolys, hipStream_t s) {
    k_sub<<<rows_grid(npolys * Ql, c.N), kBlock, 0, s>>>(a, b, c.mods(), Ql, c.N, npolys * Ql);
    SHM_HIP(hipGetLastError());
}

void launch_negate(const Context &c, uint64_t *a, size_t Ql, size_t npolys, hipStream_t s) {
    k_negate<<<rows_grid(npolys * Ql, c.N), kBlock, 0, s>>>(a, c.mods(), Ql, c.N, npolys * Ql);
    SHM_HIP(hipGetLastError());
}

void launch_mul_plain(const Context &c, uint64_t *ct, const uint64_t *pt, size_t Ql, size_t npolys, hipStream_t s) {
    k_mul_plain<<<rows_grid(npolys * Ql, c.N), kBlock, 0, s>>>(ct, pt, c.mods(), Ql, c.N, npolys * Ql);
    SHM_HIP(hipGetLastError());
}

void launch_tensor(const Context &c, const uint64_t *a, const uint64_t *b, uint64_t *out3, size_t Ql, hipStream_t s,
                   size_t a_limbs, size_t b_limbs, bool accumulate) {
    if (!a_limbs) a_limbs = Ql;
    if (!b_limbs) b_limbs = Ql;
    if (a_limbs < Ql || b_limbs < Ql) throw LogicError("tensor operand below the product level");
    k_tensor<<<rows_grid(Ql, c.N), kBlock, 0, s>>>(a, b, out3, c.mods(), Ql, c.N, a_limbs, b_limbs,
                                                   accumulate ? 1 : 0, c.primes52 ? 1 : 0);
    SHM_HIP(hipGetLastError());
}

void launch_tensor_cheb(const Context &c, const uint64_t *a, const uint64_t *b, const uint64_t *t, uint64_t cval,
                        uint64_t *out3, size_t Ql, hipStream_t s, size_t a_limbs, size_t b_limbs, size_t t_limbs) {
    if (a_limbs < Ql || b_limbs < Ql || (t && t_limbs < Ql)) throw LogicError("Chebyshev operand below the product level");
    k_tensor_cheb<<<rows_grid(Ql, c.N), kBlock, 0, s>>>(a, b, t, out3, c.mods(), Ql, c.N, a_limbs, b_limbs, t_limbs,
                                                        cval, c.primes52 ? 1 : 0);
    SHM_HIP(hipGetLastError());
}

void launch_square(const Context &c, const uint64_t *a, uint64_t *out3, size_t Ql, hipStream_t s, size_t a_limbs) {
    if (!a_limbs) a_limbs = Ql;
    if (a_limbs < Ql) throw LogicError("square operand below the product level");
    k_square<<<rows_grid(Ql, c.N), kBlock, 0, s>>>(a, out3, c.mods(), Ql, c.N, a_limbs, c.primes52 ? 1 : 0);
    SHM_HIP(hipGetLastError());
}

void launch_rescale(const Context &c, const uint64_t *in, uint64_t *out, size_t Ql, size_t npolys, hipStream_t s,
                    bool rounded) {
    if (Ql < 2) throw LogicError("end of modulus switching chain reached");
    const size_t N = c.N, L1 = Ql - 1;
    Scratch last(npolys * N * sizeof(uint64_t), s);
    Scratch delta(npolys * L1 * N * sizeof(uint64_t), s);
    launch_rescale_scratch(c, in, out, Ql, npolys, s, rounded, last.as<uint64_t>(), delta.as<uint64_t>());
}

void launch_rescale_scratch(const Context &c, const uint64_t *in, uint64_t *out, size_t Ql, size_t npolys,
                            hipStream_t s, bool rounded, uint64_t *last, uint64_t *delta) {
    if (Ql < 2) throw LogicError("end of modulus switching chain reached");
    const size_t N = c.N, L1 = Ql - 1;
    k_gather_limb<<<grid_for(npolys * N), kBlock, 0, s>>>(in, last, Ql, L1, N, npolys);
    ntt_inverse(c, last, npolys, 1, L1, s);
    const LevelConsts &L = c.level(Ql);
    // in-place: write into delta (each element is read then written by one thread), then copy back
    uint64_t *dst = (in == out) ? delta : out;
    rescale_ntt_fused(c, last, in, Ql, delta, nullptr, dst, npolys, npolys, L1, L.qlast_inv.as<uint64_t>(),
                      L.qlast_inv_s.as<uint64_t>(), s, rounded);
    if (in == out) launch_copy_words(out, delta, npolys * L1 * N, s);
    SHM_HIP(hipGetLastError());
}

void launch_gather_polys(const Context &c, const PolyPack &in, uint64_t *out, size_t words, size_t npolys,
                         hipStream_t s) {
    if (npolys == 0 || npolys > (size_t)kPolyPack) throw InvalidArgument("gather: 1..64 polynomials");
    k_gather_polys<<<rows_grid(npolys, words), kBlock, 0, s>>>(in, out, words, npolys);
    SHM_HIP(hipGetLastError());
}

void launch_permute(const Context &c, const uint64_t *in, uint64_t *out, const uint32_t *table, size_t nlimbs,
                    hipStream_t s) {
    k_permute<<<rows_grid(nlimbs, c.N), kBlock, 0, s>>>(in, out, table, c.N, nlimbs);
    SHM_HIP(hipGetLastError());
}

void launch_copy_limbs(const Context &c, const uint64_t *in, size_t in_limbs, uint64_t *out, size_t out_limbs,
                       size_t npolys, hipStream_t s) {
    k_copy_limbs<<<rows_grid(npolys * out_limbs, c.N), kBlock, 0, s>>>(in, in_limbs, out, out_limbs, c.N, npolys);
    SHM_HIP(hipGetLastError());
}

// Device-to-device copy as a kernel (16-byte accesses when both ends allow):
// it orders like every other launch on the stream, also inside a captured graph.
__global__ void k_copy_words(uint64_t *__restrict__ dst, const uint64_t *__restrict__ src, size_t words, int vec) {
    typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (vec) {
        const size_t pairs = words / 2;
        for (size_t i = t; i < pairs; i += stride) ((u64x2 *)dst)[i] = ((const u64x2 *)src)[i];
        if (t == 0 && (words & 1)) dst[words - 1] = src[words - 1];
    } else {
        for (size_t i = t; i < words; i += stride) dst[i] = src[i];
    }
}

void launch_copy_words(uint64_t *dst, const uint64_t *src, size_t words, hipStream_t s) {
    if (words == 0 || dst == src) return;
    const int vec = (((uintptr_t)dst | (uintptr_t)src) & 15) == 0 ? 1 : 0;
    k_copy_words<<<grid_for(vec ? (words + 1) / 2 : words), kBlock, 0, s>>>(dst, src, words, vec);
    SHM_HIP(hipGetLastError());
}

void launch_add_scalar(const Context &c, uint64_t *a, double rounded, size_t Ql, hipStream_t s) {
    k_add_scalar<<<rows_grid(Ql, c.N), kBlock, 0, s>>>(a, c.mods(), rounded, c.N, Ql);
    SHM_HIP(hipGetLastError());
}

void launch_pack_op(const Context &c, const PolyPack &a, const PolyPack &b, size_t npolys, double rounded, size_t Ql,
                    int op, hipStream_t s) {
    if (npolys == 0) return;
    if (npolys > (size_t)kPolyPack) throw InvalidArgument("at most 64 polynomials per batched operation");
    k_pack_op<<<rows_grid(npolys * Ql, c.N), kBlock, 0, s>>>(a, b, c.mods(), rounded, Ql, c.N, npolys * Ql, op);
    SHM_HIP(hipGetLastError());
}

void launch_mul_scalar(const Context &c, uint64_t *a, double rounded, size_t Ql, size_t npolys, hipStream_t s) {
    k_mul_scalar<<<rows_grid(npolys * Ql, c.N), kBlock, 0, s>>>(a, c.mods(), rounded, Ql, c.N, npolys * Ql);
    SHM_HIP(hipGetLastError());
}

void launch_fill_zero(uint64_t *out, size_t words, hipStream_t s) {
    if (words == 0) return;
    k_fill_zero<<<grid_for(words), kBlock, 0, s>>>(out, words);
    SHM_HIP(hipGetLastError());
}

void launch_sample_uniform(const Context &c, uint64_t seed, uint64_t *out, size_t Ql, bool withP, hipStream_t s) {
    const size_t limbs = Ql + (withP ? c.nP : 0), total = limbs * c.N;
    k_sample_uniform<<<grid_for(total), kBlock, 0, s>>>(seed, out, c.mods(), Ql, c.nQ, c.N, total);
    SHM_HIP(hipGetLastError());
}

void launch_sample_gaussian_rns(const Context &c, uint64_t seed, uint64_t *out, size_t Ql, bool withP,
                                hipStream_t s) {
    const size_t limbs = Ql + (withP ? c.nP : 0), total = limbs * c.N;
    k_sample_gaussian<<<grid_for(total), kBlock, 0, s>>>(seed, out, c.mods(), Ql, c.nQ, c.N, total);
    SHM_HIP(hipGetLastError());
}

void launch_signed_to_rns(const Context &c, const int8_t *coeffs, uint64_t *out, size_t Ql, bool withP,
                          hipStream_t s) {
    const size_t limbs = Ql + (withP ? c.nP : 0), total = limbs * c.N;
    k_signed_to_rns<<<grid_for(total), kBlock, 0, s>>>(coeffs, out, c.mods(), Ql, c.nQ, c.N, total);
    SHM_HIP(hipGetLastError());
}

void launch_ks_key_digit(const Context &c, const uint64_t *sk, const uint64_t *newkey, uint64_t *b,
                         const uint64_t *a, const uint64_t *e, size_t digit, hipStream_t s) {
    const size_t total = c.nQP() * c.N;
    k_ks_key_digit<<<grid_for(total), kBlock, 0, s>>>(sk, newkey, b, a, e, c.mods(), c.pmod.as<uint64_t>(),
                                                      c.pmod_s.as<uint64_t>(), digit * c.nP, (digit + 1) * c.nP,
                                                      c.N, total);
    SHM_HIP(hipGetLastError());
}

void launch_encrypt_combine(const Context &c, const uint64_t *sk, const uint64_t *a, const uint64_t *e,
                            const uint64_t *pt, uint64_t *c0, size_t Ql, hipStream_t s) {
    const size_t total = Ql * c.N;
    k_encrypt_combine<<<grid_for(total), kBlock, 0, s>>>(sk, a, e, pt, c0, c.mods(), c.N, total);
    SHM_HIP(hipGetLastError());
}

void launch_decrypt(const Context &c, const uint64_t *ct, size_t size, const uint64_t *sk, uint64_t *pt, size_t Ql,
                    hipStream_t s) {
    k_decrypt<<<grid_for(Ql * c.N), kBlock, 0, s>>>(ct, size, sk, pt, c.mods(), c.N, Ql * c.N);
    SHM_HIP(hipGetLastError());
}

void launch_mul_elementwise(const Context &c, const uint64_t *a, const uint64_t *b, uint64_t *out, size_t nlimbs,
                            bool withP, size_t Ql, hipStream_t s) {
    (void)withP;
    const size_t total = nlimbs * c.N;
    k_mul_elementwise<<<grid_for(total), kBlock, 0, s>>>(a, b, out, c.mods(), Ql, c.nQ, c.N, total);
    SHM_HIP(hipGetLastError());
}

}  // namespace shm
