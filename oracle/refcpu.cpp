/*
 * refcpu — CPU restatement of the ShMOAI / PhantomFHE CKKS limb arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle). It exists to check the HIP engine in
 * shmoai_amd/ and to time a CPU baseline; the product never links it.
 *
 * Citations are reference file:line (Phantom sources are under
 * thirdparty/phantom-fhe/). Values are canonical residues in [0, q) wherever
 * the reference produces canonical residues; the reference's lazy [0,2q) /
 * [0,4q) intermediates are not modelled because every reference kernel ends
 * with a final reduction (fntt_2d.cu:188-196, intt_2d.cu:194-203), so the
 * canonical result is fully determined by the mathematics restated here.
 *
 * Samplers: the reference's Salsa20 streams (prng.cu:17-240) and its unseeded
 * Hamming-weight shuffle (secretkey.cu:389-390) cannot be reproduced, so key
 * material comes from injected, seeded samplers that the GPU engine restates
 * bit-for-bit: SplitMix64 (Steele, Lea, Flood 2014) counters; uniform residues
 * from a 128-bit draw reduced mod q; a discrete Gaussian sigma = 3.2 by
 * cumulative-distribution table (|e| <= 19); ternary secrets of exact Hamming
 * weight by a partial Fisher-Yates shuffle.
 */
#include "refcpu.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <functional>
#include <cstring>
#include <map>
#include <stdexcept>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;

namespace {

int g_threads = 1;

inline uint64_t mulmod(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)(((u128)a * b) % q); }
inline uint64_t addmod(uint64_t a, uint64_t b, uint64_t q) {
    uint64_t s = a + b;
    return s >= q ? s - q : s;
}
inline uint64_t submod(uint64_t a, uint64_t b, uint64_t q) { return a >= b ? a - b : a + q - b; }
inline uint64_t negmod(uint64_t a, uint64_t q) { return a == 0 ? 0 : q - a; }
inline uint64_t shoup(uint64_t w, uint64_t q) { return (uint64_t)(((u128)w << 64) / q); }
inline uint64_t mul_shoup(uint64_t x, uint64_t w, uint64_t ws, uint64_t q) {
    uint64_t hi = (uint64_t)(((u128)x * ws) >> 64);
    uint64_t r = x * w - hi * q;
    return r >= q ? r - q : r;
}
uint64_t powmod(uint64_t b, uint64_t e, uint64_t q) {
    uint64_t r = 1 % q;
    b %= q;
    while (e) {
        if (e & 1) r = mulmod(r, b, q);
        b = mulmod(b, b, q);
        e >>= 1;
    }
    return r;
}
uint64_t invmod(uint64_t a, uint64_t q) { return powmod(a, q - 2, q); } /* q prime */

uint32_t brv(uint32_t x, int bits) {
    uint32_t r = 0;
    for (int i = 0; i < bits; i++) r |= ((x >> i) & 1u) << (bits - 1 - i);
    return r;
}

/* Deterministic Miller-Rabin (bases 2..37 are exact for 64-bit inputs). The
 * reference uses randomised rounds (util/numth.cu:131-207); both decide
 * primality exactly for the candidates involved. */
bool is_prime_u64(uint64_t n) {
    if (n < 2) return false;
    static const uint64_t small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    for (uint64_t p : small) {
        if (n % p == 0) return n == p;
    }
    uint64_t d = n - 1;
    int r = 0;
    while ((d & 1) == 0) {
        d >>= 1;
        r++;
    }
    for (uint64_t a : small) {
        uint64_t x = powmod(a, d, n);
        if (x == 1 || x == n - 1) continue;
        bool comp = true;
        for (int i = 1; i < r; i++) {
            x = mulmod(x, x, n);
            if (x == n - 1) {
                comp = false;
                break;
            }
        }
        if (comp) return false;
    }
    return true;
}

/* util/numth.cu:209-235 get_primes: descending from 2^b - 2N + 1 in steps of 2N. */
std::vector<uint64_t> get_primes(size_t N, int bit_size, size_t count) {
    std::vector<uint64_t> out;
    uint64_t factor = 2 * (uint64_t)N;
    uint64_t value = (uint64_t(1) << bit_size) - factor + 1;
    uint64_t lower = uint64_t(1) << (bit_size - 1);
    while (count > 0 && value > lower) {
        if (is_prime_u64(value)) {
            out.push_back(value);
            count--;
        }
        value -= factor;
    }
    if (count > 0) throw std::logic_error("failed to find enough qualifying primes");
    return out;
}

/* util/numth.cu:278-333 try_minimal_primitive_root: the smallest primitive
 * 2N-th root of unity (the reference scans all odd powers of a random one). */
uint64_t minimal_primitive_root(uint64_t degree, uint64_t q) {
    uint64_t quot = (q - 1) / degree;
    uint64_t root = 0;
    for (uint64_t x = 2; x < q; x++) {
        uint64_t r = powmod(x, quot, q);
        if (powmod(r, degree >> 1, q) == q - 1) {
            root = r;
            break;
        }
    }
    uint64_t g2 = mulmod(root, root, q), cur = root, best = root;
    for (uint64_t i = 0; i < degree; i++) {
        if (cur < best) best = cur;
        cur = mulmod(cur, g2, q);
    }
    return best;
}

const uint64_t kCDT[19] = {
    0x0ff52b40a5917f1dULL, 0x2e5a25d4bf0e400eULL, 0x489ae26955b04bd6ULL, 0x5d2bc20f621bf185ULL,
    0x6bc8694c3cc80ff4ULL, 0x7532d89ac6ba7dceULL, 0x7ab396cb74436798ULL, 0x7d9e4e916643eb07ULL,
    0x7f05495819eb2051ULL, 0x7fa1ce9c0039a957ULL, 0x7fdfb3f212e8c4e8ULL, 0x7ff5e6f9d2314fccULL,
    0x7ffd1f97bc4406a2ULL, 0x7fff40fa0088d11dULL, 0x7fffd2e835e1c57dULL, 0x7ffff6524386ff1eULL,
    0x7ffffe1db4769da5ULL, 0x7fffffac0a1dcb08ULL, 0x7ffffff428673853ULL};

uint64_t sm64(uint64_t key, uint64_t ctr) {
    uint64_t z = key + (ctr + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

}  // namespace

struct RefCtx {
    int logN;
    size_t N, nQ, nP;
    std::vector<uint64_t> q, psi, ninv, ninv_s;
    std::vector<std::vector<uint64_t>> tw, tws, itw, itws;
    /* encoder: ckks.cu:30-87 */
    std::vector<std::complex<double>> roots; /* m = 2N powers */
    std::vector<uint32_t> rot_group;         /* N/4 powers of 5 mod 2N */

    size_t nQP() const { return nQ + nP; }
    uint32_t pidx(size_t k, size_t Ql) const { return (uint32_t)(k < Ql ? k : nQ + (k - Ql)); }
};

namespace {

void ntt_limb(const RefCtx &c, uint64_t *a, size_t pi) {
    /* fntt_2d.cu:9-197 + butterfly.cuh:10-23: Cooley-Tukey, stage m uses
     * tw[m + i] where tw[brv(k)] = psi^k (util/ntt.cu:24-33). */
    const size_t N = c.N;
    const uint64_t q = c.q[pi];
    const uint64_t *tw = c.tw[pi].data(), *ts = c.tws[pi].data();
    for (size_t j = 0; j < N; j++) a[j] %= q;
    size_t t = N;
    for (size_t m = 1; m < N; m <<= 1) {
        t >>= 1;
        for (size_t i = 0; i < m; i++) {
            const uint64_t w = tw[m + i], ws = ts[m + i];
            uint64_t *x = a + 2 * i * t, *y = x + t;
            for (size_t j = 0; j < t; j++) {
                uint64_t u = x[j], v = mul_shoup(y[j], w, ws, q);
                x[j] = addmod(u, v, q);
                y[j] = submod(u, v, q);
            }
        }
    }
}

void intt_limb(const RefCtx &c, uint64_t *a, size_t pi) {
    /* intt_2d.cu:9-207 + butterfly.cuh:25-37: Gentleman-Sande with the
     * inverse table itw[brv(k)] = psi^-k, then x n^-1 (folded into the last
     * stage by the reference, util/ntt.cu:52-55). */
    const size_t N = c.N;
    const uint64_t q = c.q[pi];
    const uint64_t *tw = c.itw[pi].data(), *ts = c.itws[pi].data();
    for (size_t j = 0; j < N; j++) a[j] %= q;
    size_t t = 1;
    for (size_t m = N >> 1; m >= 1; m >>= 1) {
        for (size_t i = 0; i < m; i++) {
            const uint64_t w = tw[m + i], ws = ts[m + i];
            uint64_t *x = a + 2 * i * t, *y = x + t;
            for (size_t j = 0; j < t; j++) {
                uint64_t u = x[j], v = y[j];
                x[j] = addmod(u, v, q);
                y[j] = mul_shoup(submod(u, v, q), w, ws, q);
            }
        }
        t <<= 1;
    }
    for (size_t j = 0; j < N; j++) a[j] = mul_shoup(a[j], c.ninv[pi], c.ninv_s[pi], q);
}

void ntt_limbs(const RefCtx &c, uint64_t *data, size_t nlimbs, const uint32_t *pidx) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (long k = 0; k < (long)nlimbs; k++) ntt_limb(c, data + k * c.N, pidx[k]);
}
void intt_limbs(const RefCtx &c, uint64_t *data, size_t nlimbs, const uint32_t *pidx) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (long k = 0; k < (long)nlimbs; k++) intt_limb(c, data + k * c.N, pidx[k]);
}

std::vector<uint32_t> idx_Ql(const RefCtx &c, size_t Ql, bool withP) {
    std::vector<uint32_t> v;
    for (size_t k = 0; k < Ql; k++) v.push_back((uint32_t)k);
    if (withP)
        for (size_t k = 0; k < c.nP; k++) v.push_back((uint32_t)(c.nQ + k));
    return v;
}

/* signed coefficient -> residue mod q */
inline uint64_t from_signed(int64_t v, uint64_t q) {
    if (v >= 0) return (uint64_t)v % q;
    uint64_t m = (uint64_t)(-(v + 1)) + 1; /* |v| without overflow */
    m %= q;
    return m == 0 ? 0 : q - m;
}

void galois_table(const RefCtx &c, uint32_t elt, std::vector<uint32_t> &table) {
    /* galois.cuh (PhantomGaloisTool ctor): for i in [N, 2N),
     * table[i-N] = brv_logN(((elt * brv_{logN+1}(i)) >> 1) & (N-1)). */
    table.resize(c.N);
    for (size_t i = c.N; i < 2 * c.N; i++) {
        uint32_t r = brv((uint32_t)i, c.logN + 1);
        uint64_t raw = ((uint64_t)elt * r) >> 1;
        raw &= (uint64_t)(c.N - 1);
        table[i - c.N] = brv((uint32_t)raw, c.logN);
    }
}

/* (prod_{k in S, k != i} q_k) mod p */
uint64_t punctured_prod_mod(const RefCtx &c, const std::vector<uint32_t> &S, size_t i, uint64_t p) {
    uint64_t r = 1 % p;
    for (size_t k = 0; k < S.size(); k++)
        if (k != i) r = mulmod(r, c.q[S[k]] % p, p);
    return r;
}

}  // namespace

extern "C" {

void ref_set_threads(int n) { g_threads = n < 1 ? 1 : n; }

int ref_is_prime(uint64_t v) { return is_prime_u64(v) ? 1 : 0; }

void ref_get_primes(size_t N, int bit_size, size_t count, uint64_t *out) {
    auto v = get_primes(N, bit_size, count);
    std::copy(v.begin(), v.end(), out);
}

RefCtx *ref_ctx_new(int logN, int n_bits, const int *bit_sizes, int n_special) {
    RefCtx *c = new RefCtx();
    c->logN = logN;
    c->N = size_t(1) << logN;
    c->nP = (size_t)n_special;
    c->nQ = (size_t)n_bits - c->nP;
    /* modulus.cu:79-108 CoeffModulus::Create over the whole list (data and
     * special primes together); each bit-size class is filled from the back
     * of its get_primes list. */
    std::map<int, size_t> count;
    for (int i = 0; i < n_bits; i++) count[bit_sizes[i]]++;
    std::map<int, std::vector<uint64_t>> table;
    for (auto &kv : count) table[kv.first] = get_primes(c->N, kv.first, kv.second);
    for (int i = 0; i < n_bits; i++) {
        auto &v = table[bit_sizes[i]];
        c->q.push_back(v.back());
        v.pop_back();
    }
    const size_t N = c->N, L = c->q.size();
    c->psi.resize(L);
    c->ninv.resize(L);
    c->ninv_s.resize(L);
    c->tw.assign(L, std::vector<uint64_t>(N));
    c->tws.assign(L, std::vector<uint64_t>(N));
    c->itw.assign(L, std::vector<uint64_t>(N));
    c->itws.assign(L, std::vector<uint64_t>(N));
    for (size_t p = 0; p < L; p++) {
        const uint64_t q = c->q[p];
        /* util/ntt.cu:11-56 */
        uint64_t psi = minimal_primitive_root(2 * N, q);
        uint64_t ipsi = invmod(psi, q);
        c->psi[p] = psi;
        uint64_t pw = 1, ipw = 1;
        for (size_t i = 0; i < N; i++) {
            uint32_t r = brv((uint32_t)i, logN);
            c->tw[p][r] = pw;
            c->itw[p][r] = ipw;
            pw = mulmod(pw, psi, q);
            ipw = mulmod(ipw, ipsi, q);
        }
        for (size_t i = 0; i < N; i++) {
            c->tws[p][i] = shoup(c->tw[p][i], q);
            c->itws[p][i] = shoup(c->itw[p][i], q);
        }
        c->ninv[p] = invmod(N % q, q);
        c->ninv_s[p] = shoup(c->ninv[p], q);
    }
    /* ckks.cu:30-87 + fft.cu:11-40 */
    const size_t m = 2 * N;
    const double PI_ = 3.1415926535897932384626433832795028842;
    std::vector<std::complex<double>> base(m / 8 + 1);
    for (size_t i = 0; i <= m / 8; i++)
        base[i] = std::complex<double>(std::cos(2 * PI_ * (double)i / (double)m),
                                       std::sin(2 * PI_ * (double)i / (double)m));
    std::function<std::complex<double>(size_t)> get_root = [&](size_t index) -> std::complex<double> {
        index &= m - 1;
        if (index <= m / 8) return base[index];
        if (index <= m / 4) {
            auto a = base[m / 4 - index];
            return std::complex<double>(a.imag(), a.real());
        }
        if (index <= m / 2) {
            auto a = get_root(m / 2 - index);
            return std::complex<double>(0.0 - a.real(), 0.0 - (-a.imag()));
        }
        if (index <= 3 * m / 4) {
            auto a = get_root(index - m / 2);
            return std::complex<double>(0.0 - a.real(), 0.0 - a.imag());
        }
        auto a = get_root(m - index);
        return std::complex<double>(a.real(), -a.imag());
    };
    c->roots.resize(m);
    for (size_t i = 0; i < m; i++) c->roots[i] = get_root(i);
    c->rot_group.resize(N / 4);
    uint32_t pos = 1;
    for (size_t i = 0; i < N / 4; i++) {
        c->rot_group[i] = pos;
        pos = (uint32_t)(((uint64_t)pos * 5) & (m - 1));
    }
    return c;
}

void ref_ctx_free(RefCtx *c) { delete c; }
size_t ref_ctx_nq(const RefCtx *c) { return c->nQ; }
size_t ref_ctx_np(const RefCtx *c) { return c->nP; }
void ref_ctx_primes(const RefCtx *c, uint64_t *out) { std::copy(c->q.begin(), c->q.end(), out); }
void ref_ctx_psi(const RefCtx *c, uint64_t *out) { std::copy(c->psi.begin(), c->psi.end(), out); }
void ref_ctx_tables(const RefCtx *c, size_t prime, uint64_t *tw, uint64_t *itw) {
    std::copy(c->tw[prime].begin(), c->tw[prime].end(), tw);
    std::copy(c->itw[prime].begin(), c->itw[prime].end(), itw);
}

void ref_ntt(const RefCtx *c, uint64_t *data, size_t nlimbs, const uint32_t *prime_idx) {
    ntt_limbs(*c, data, nlimbs, prime_idx);
}
void ref_intt(const RefCtx *c, uint64_t *data, size_t nlimbs, const uint32_t *prime_idx) {
    intt_limbs(*c, data, nlimbs, prime_idx);
}

/* polymath.cu:41-125 add/sub/negate; polymath.cu:161-185 multiply_rns_poly. */
void ref_add(const RefCtx *c, const uint64_t *a, const uint64_t *b, uint64_t *out, size_t Ql, size_t npolys) {
    for (size_t p = 0; p < npolys; p++)
        for (size_t k = 0; k < Ql; k++) {
            const uint64_t q = c->q[k];
            size_t o = (p * Ql + k) * c->N;
            for (size_t j = 0; j < c->N; j++) out[o + j] = addmod(a[o + j], b[o + j], q);
        }
}
void ref_sub(const RefCtx *c, const uint64_t *a, const uint64_t *b, uint64_t *out, size_t Ql, size_t npolys) {
    for (size_t p = 0; p < npolys; p++)
        for (size_t k = 0; k < Ql; k++) {
            const uint64_t q = c->q[k];
            size_t o = (p * Ql + k) * c->N;
            for (size_t j = 0; j < c->N; j++) out[o + j] = submod(a[o + j], b[o + j], q);
        }
}
void ref_negate(const RefCtx *c, const uint64_t *a, uint64_t *out, size_t Ql, size_t npolys) {
    for (size_t p = 0; p < npolys; p++)
        for (size_t k = 0; k < Ql; k++) {
            const uint64_t q = c->q[k];
            size_t o = (p * Ql + k) * c->N;
            for (size_t j = 0; j < c->N; j++) out[o + j] = negmod(a[o + j], q);
        }
}
/* evaluate.cu:1237-1266 multiply_plain_ntt: c_k <- c_k (.) pt per limb. */
void ref_multiply_plain(const RefCtx *c, uint64_t *ct, const uint64_t *pt, size_t Ql, size_t npolys) {
    for (size_t p = 0; p < npolys; p++)
        for (size_t k = 0; k < Ql; k++) {
            const uint64_t q = c->q[k];
            uint64_t *x = ct + (p * Ql + k) * c->N;
            const uint64_t *y = pt + k * c->N;
            for (size_t j = 0; j < c->N; j++) x[j] = mulmod(x[j], y[j], q);
        }
}
/* polymath.cu:468-503 tensor_prod_2x2_rns_poly (Karatsuba middle term). */
void ref_tensor(const RefCtx *c, const uint64_t *a, const uint64_t *b, uint64_t *out3, size_t Ql) {
    const size_t N = c->N, S = Ql * N;
    for (size_t k = 0; k < Ql; k++) {
        const uint64_t q = c->q[k];
        for (size_t j = 0; j < N; j++) {
            size_t i = k * N + j;
            uint64_t a0 = a[i], a1 = a[i + S], b0 = b[i], b1 = b[i + S];
            uint64_t d0 = mulmod(a0, b0, q), d2 = mulmod(a1, b1, q);
            uint64_t d1 = mulmod(addmod(a0, a1, q), addmod(b0, b1, q), q);
            d1 = submod(submod(d1, d0, q), d2, q);
            out3[i] = d0;
            out3[i + S] = d1;
            out3[i + 2 * S] = d2;
        }
    }
}

/* rns.cu:1125-1181 divide_and_round_q_last_ntt (floor division by q_last):
 * iNTT(last limb) -> [.]_{q_j} -> NTT -> (c_j - .) * q_last^-1. */
void ref_rescale(const RefCtx *c, const uint64_t *in, uint64_t *out, size_t Ql, size_t npolys) {
    const size_t N = c->N, L1 = Ql - 1;
    const uint64_t ql = c->q[Ql - 1];
    std::vector<uint64_t> last(N), tmp(L1 * N);
    for (size_t p = 0; p < npolys; p++) {
        std::copy(in + (p * Ql + L1) * N, in + (p * Ql + Ql) * N, last.begin());
        intt_limb(*c, last.data(), L1);
        for (size_t k = 0; k < L1; k++)
            for (size_t j = 0; j < N; j++) tmp[k * N + j] = last[j] % c->q[k];
        auto idx = idx_Ql(*c, L1, false);
        ntt_limbs(*c, tmp.data(), L1, idx.data());
        for (size_t k = 0; k < L1; k++) {
            const uint64_t q = c->q[k], inv = invmod(ql % q, q);
            for (size_t j = 0; j < N; j++)
                out[(p * L1 + k) * N + j] = mulmod(submod(in[(p * Ql + k) * N + j], tmp[k * N + j], q), inv, q);
        }
    }
}
/* evaluate.cu:1445-1561 mod_switch_drop_to_next: drop the last limb. */
void ref_mod_switch_drop(const RefCtx *c, const uint64_t *in, uint64_t *out, size_t Ql, size_t npolys) {
    const size_t N = c->N;
    for (size_t p = 0; p < npolys; p++)
        std::memcpy(out + p * (Ql - 1) * N, in + p * Ql * N, (Ql - 1) * N * sizeof(uint64_t));
}

/* galois.cu:47-90 get_elt_from_step (generator 5; step<0 -> N/2 - |step|;
 * step 0 -> conjugation 2N-1). */
uint32_t ref_galois_elt_from_step(const RefCtx *c, int step) {
    uint64_t m = 2 * (uint64_t)c->N;
    if (step == 0) return (uint32_t)(m - 1);
    uint32_t pos = (uint32_t)std::abs(step);
    if (pos >= (c->N >> 1)) throw std::invalid_argument("step count too large");
    int s = step < 0 ? (int)(c->N >> 1) - (int)pos : (int)pos;
    uint64_t e = 1;
    while (s--) e = (e * 5) & (m - 1);
    return (uint32_t)e;
}
void ref_galois_table(const RefCtx *c, uint32_t elt, uint32_t *table) {
    std::vector<uint32_t> t;
    galois_table(*c, elt, t);
    std::copy(t.begin(), t.end(), table);
}
/* galois.cu:8-20 apply_galois_ntt_permutation: dst[i] = src[perm[i]] per limb. */
void ref_apply_galois_ntt(const RefCtx *c, uint32_t elt, const uint64_t *in, uint64_t *out, size_t nlimbs) {
    std::vector<uint32_t> t;
    galois_table(*c, elt, t);
    for (size_t k = 0; k < nlimbs; k++)
        for (size_t j = 0; j < c->N; j++) out[k * c->N + j] = in[k * c->N + t[j]];
}

size_t ref_beta(const RefCtx *c, size_t Ql) { return (Ql + c->nP - 1) / c->nP; }

/* rns_bconv.cu:530-634 DRNSTool::modup with rns.cu:131-186 constants.
 * out: beta x QlP x N, NTT form. Own digit limbs are copied from the NTT
 * input; complement limbs are the approximate fast base conversion
 * sum_i [x_i * qhat_i^-1]_{q_i} * [qhat_i]_p (alpha = 1: [x]_p) followed by
 * NTT over the complement primes (ntt_modup.cu:610-657). */
void ref_modup(const RefCtx *c, const uint64_t *c2, size_t Ql, uint64_t *out) {
    const size_t N = c->N, alpha = c->nP, QlP = Ql + c->nP, beta = ref_beta(c, Ql);
    std::vector<uint64_t> t(c2, c2 + Ql * N);
    auto idxQ = idx_Ql(*c, Ql, false);
    intt_limbs(*c, t.data(), Ql, idxQ.data());
    for (size_t g = 0; g < beta; g++) {
        const size_t start = alpha * g, size = std::min(alpha, Ql - start);
        std::vector<uint32_t> digit;
        for (size_t i = 0; i < size; i++) digit.push_back((uint32_t)(start + i));
        std::vector<uint64_t> tt(size * N);
        for (size_t i = 0; i < size; i++) {
            const uint64_t q = c->q[start + i];
            if (alpha == 1) {
                std::copy(t.begin() + (start + i) * N, t.begin() + (start + i + 1) * N, tt.begin() + i * N);
            } else {
                /* partQlHatInv (rns.cu:158-161) */
                uint64_t hinv = invmod(punctured_prod_mod(*c, digit, i, q), q);
                for (size_t j = 0; j < N; j++) tt[i * N + j] = mulmod(t[(start + i) * N + j], hinv, q);
            }
        }
        uint64_t *d = out + g * QlP * N;
        std::vector<uint32_t> comp_idx;
        std::vector<size_t> comp_k;
        for (size_t k = 0; k < QlP; k++) {
            if (k >= start && k < start + size) {
                std::memcpy(d + k * N, c2 + k * N, N * sizeof(uint64_t));
                continue;
            }
            const uint32_t pi = c->pidx(k, Ql);
            const uint64_t p = c->q[pi];
            uint64_t *dk = d + k * N;
            if (alpha == 1) {
                for (size_t j = 0; j < N; j++) dk[j] = tt[j] % p;
            } else {
                std::vector<uint64_t> hp(size);
                for (size_t i = 0; i < size; i++) hp[i] = punctured_prod_mod(*c, digit, i, p);
                for (size_t j = 0; j < N; j++) {
                    u128 acc = 0;
                    for (size_t i = 0; i < size; i++) acc += (u128)tt[i * N + j] * hp[i];
                    dk[j] = (uint64_t)(acc % p);
                }
            }
            comp_idx.push_back(pi);
            comp_k.push_back(k);
        }
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (long z = 0; z < (long)comp_k.size(); z++) ntt_limb(*c, d + comp_k[z] * N, comp_idx[z]);
    }
}

/* eval_key_switch.cu:12-60 key_switch_inner_prod_c2_and_evk: over QlP limbs,
 * cx0 = sum_g d_g * key_g.b, cx1 = sum_g d_g * key_g.a (key limb for special
 * prime k >= Ql is size_Q + (k - Ql)). cx: 2 x QlP x N. */
void ref_ks_inner(const RefCtx *c, const uint64_t *modup, const uint64_t *key, size_t Ql, uint64_t *cx) {
    const size_t N = c->N, QlP = Ql + c->nP, beta = ref_beta(c, Ql), QP = c->nQP();
    for (size_t k = 0; k < QlP; k++) {
        const uint32_t pi = c->pidx(k, Ql);
        const uint64_t q = c->q[pi];
        for (size_t j = 0; j < N; j++) {
            u128 a0 = 0, a1 = 0;
            for (size_t g = 0; g < beta; g++) {
                uint64_t dv = modup[(g * QlP + k) * N + j];
                const uint64_t *kg = key + g * 2 * QP * N;
                a0 += (u128)dv * kg[pi * N + j];
                a1 += (u128)dv * kg[(QP + pi) * N + j];
            }
            cx[k * N + j] = (uint64_t)(a0 % q);
            cx[(QlP + k) * N + j] = (uint64_t)(a1 % q);
        }
    }
}

/* rns_bconv.cu:776-828 moddown_from_NTT (CKKS) + ntt_moddown.cu:106-261:
 * iNTT(P limbs) -> delta = fastbconv(P -> Ql) (alpha=1: [c_P]_q) -> NTT ->
 * out = (cx - delta) * P^-1 mod q. cx (QlP limbs) is modified (P limbs). */
void ref_moddown(const RefCtx *c, uint64_t *cx, size_t Ql, uint64_t *out) {
    const size_t N = c->N, nP = c->nP;
    std::vector<uint32_t> pP;
    for (size_t k = 0; k < nP; k++) pP.push_back((uint32_t)(c->nQ + k));
    intt_limbs(*c, cx + Ql * N, nP, pP.data());
    std::vector<uint64_t> delta(Ql * N);
    std::vector<uint64_t> tP(nP * N);
    for (size_t jj = 0; jj < nP; jj++) {
        const uint64_t p = c->q[c->nQ + jj];
        if (nP == 1) {
            std::memcpy(tP.data(), cx + Ql * N, N * sizeof(uint64_t));
        } else {
            uint64_t hinv = invmod(punctured_prod_mod(*c, pP, jj, p), p);
            for (size_t j = 0; j < N; j++) tP[jj * N + j] = mulmod(cx[(Ql + jj) * N + j], hinv, p);
        }
    }
    for (size_t i = 0; i < Ql; i++) {
        const uint64_t q = c->q[i];
        if (nP == 1) {
            for (size_t j = 0; j < N; j++) delta[i * N + j] = tP[j] % q;
        } else {
            std::vector<uint64_t> hq(nP);
            for (size_t jj = 0; jj < nP; jj++) hq[jj] = punctured_prod_mod(*c, pP, jj, q);
            for (size_t j = 0; j < N; j++) {
                u128 acc = 0;
                for (size_t jj = 0; jj < nP; jj++) acc += (u128)tP[jj * N + j] * hq[jj];
                delta[i * N + j] = (uint64_t)(acc % q);
            }
        }
    }
    auto idxQ = idx_Ql(*c, Ql, false);
    ntt_limbs(*c, delta.data(), Ql, idxQ.data());
    for (size_t i = 0; i < Ql; i++) {
        const uint64_t q = c->q[i];
        uint64_t P = 1;
        for (size_t jj = 0; jj < nP; jj++) P = mulmod(P, c->q[c->nQ + jj] % q, q);
        const uint64_t Pinv = invmod(P, q);
        for (size_t j = 0; j < N; j++)
            out[i * N + j] = mulmod(submod(cx[i * N + j], delta[i * N + j], q), Pinv, q);
    }
}

/* eval_key_switch.cu:91-178 keyswitch_inplace (without the final add). */
void ref_keyswitch(const RefCtx *c, const uint64_t *c2, const uint64_t *key, size_t Ql, uint64_t *ks0,
                   uint64_t *ks1) {
    const size_t N = c->N, QlP = Ql + c->nP, beta = ref_beta(c, Ql);
    std::vector<uint64_t> mu(beta * QlP * N), cx(2 * QlP * N);
    ref_modup(c, c2, Ql, mu.data());
    ref_ks_inner(c, mu.data(), key, Ql, cx.data());
    ref_moddown(c, cx.data(), Ql, ks0);
    ref_moddown(c, cx.data() + QlP * N, Ql, ks1);
}

/* evaluate.cu:1586-1647 apply_galois_inplace: (sigma(c0) + ks0, ks1) with
 * ks = keyswitch(sigma(c1)). */
void ref_rotate(const RefCtx *c, const uint64_t *ct, size_t Ql, uint32_t elt, const uint64_t *key, uint64_t *out) {
    const size_t N = c->N, S = Ql * N;
    std::vector<uint64_t> p0(S), p1(S), k0(S), k1(S);
    ref_apply_galois_ntt(c, elt, ct, p0.data(), Ql);
    ref_apply_galois_ntt(c, elt, ct + S, p1.data(), Ql);
    ref_keyswitch(c, p1.data(), key, Ql, k0.data(), k1.data());
    for (size_t k = 0; k < Ql; k++)
        for (size_t j = 0; j < N; j++) {
            size_t i = k * N + j;
            out[i] = addmod(p0[i], k0[i], c->q[k]);
            out[S + i] = k1[i];
        }
}

/* evaluate.cu:1358-1390 relinearize_inplace: (d0 + ks0, d1 + ks1), ks = keyswitch(d2). */
void ref_relinearize(const RefCtx *c, const uint64_t *ct3, size_t Ql, const uint64_t *key, uint64_t *out) {
    const size_t N = c->N, S = Ql * N;
    std::vector<uint64_t> k0(S), k1(S);
    ref_keyswitch(c, ct3 + 2 * S, key, Ql, k0.data(), k1.data());
    for (size_t k = 0; k < Ql; k++)
        for (size_t j = 0; j < N; j++) {
            size_t i = k * N + j;
            out[i] = addmod(ct3[i], k0[i], c->q[k]);
            out[S + i] = addmod(ct3[S + i], k1[i], c->q[k]);
        }
}

uint64_t ref_splitmix64(uint64_t key, uint64_t ctr) { return sm64(key, ctr); }

void ref_sample_uniform(const RefCtx *c, uint64_t seed, size_t nlimbs, const uint32_t *prime_idx, uint64_t *out) {
    const size_t N = c->N;
    for (size_t k = 0; k < nlimbs; k++) {
        const uint32_t pi = prime_idx[k];
        const uint64_t q = c->q[pi];
        for (size_t j = 0; j < N; j++) {
            uint64_t ctr = 2 * ((uint64_t)pi * N + j);
            u128 v = ((u128)sm64(seed, ctr) << 64) | sm64(seed, ctr + 1);
            out[k * N + j] = (uint64_t)(v % q);
        }
    }
}

void ref_sample_gaussian(const RefCtx *c, uint64_t seed, int64_t *out) {
    for (size_t j = 0; j < c->N; j++) {
        uint64_t u = sm64(seed, j), v = u >> 1;
        int64_t mag = 0;
        while (mag < 19 && v >= kCDT[mag]) mag++;
        out[j] = (u & 1) ? -mag : mag;
    }
}

void ref_sample_ternary(const RefCtx *c, uint64_t seed, size_t hw, int64_t *out) {
    const size_t N = c->N;
    std::vector<uint32_t> perm(N);
    for (size_t i = 0; i < N; i++) perm[i] = (uint32_t)i;
    for (size_t i = 0; i < N; i++) out[i] = 0;
    for (size_t i = 0; i < hw; i++) {
        size_t j = i + (size_t)(sm64(seed, i) % (uint64_t)(N - i));
        std::swap(perm[i], perm[j]);
        out[perm[i]] = (sm64(seed, N + i) & 1) ? -1 : 1;
    }
}

/* secretkey.cu:431-522 gen_secretkey: s in NTT form over all of QP. */
void ref_sk_ntt(const RefCtx *c, const int64_t *coeffs, uint64_t *sk_qp) {
    const size_t N = c->N, L = c->nQP();
    for (size_t k = 0; k < L; k++)
        for (size_t j = 0; j < N; j++) sk_qp[k * N + j] = from_signed(coeffs[j], c->q[k]);
    auto idx = idx_Ql(*c, c->nQ, true);
    ntt_limbs(*c, sk_qp, L, idx.data());
}

/* secretkey.cu:328-373 generate_one_kswitch_key + 260-326
 * encrypt_zero_symmetric: digit g = (-(a_g s + e_g) + P [newkey]_{digit g}, a_g)
 * at QP, dnum = size_Q / size_P digits. Digit g draws a from splitmix(seed,2g)
 * and e from splitmix(seed,2g+1). key: dnum x [2][QP][N]. */
void ref_gen_kswitch_key(const RefCtx *c, const uint64_t *sk_qp, const uint64_t *newkey_qp, uint64_t seed,
                         uint64_t *key) {
    const size_t N = c->N, QP = c->nQP(), alpha = c->nP, dnum = c->nQ / c->nP;
    auto idx = idx_Ql(*c, c->nQ, true);
    std::vector<int64_t> e(N);
    std::vector<uint64_t> er(QP * N);
    for (size_t g = 0; g < dnum; g++) {
        uint64_t *b = key + g * 2 * QP * N, *a = b + QP * N;
        ref_sample_uniform(c, sm64(seed, 2 * g), QP, idx.data(), a);
        ref_sample_gaussian(c, sm64(seed, 2 * g + 1), e.data());
        for (size_t k = 0; k < QP; k++)
            for (size_t j = 0; j < N; j++) er[k * N + j] = from_signed(e[j], c->q[k]);
        ntt_limbs(*c, er.data(), QP, idx.data());
        for (size_t k = 0; k < QP; k++) {
            const uint64_t q = c->q[k];
            for (size_t j = 0; j < N; j++) {
                size_t i = k * N + j;
                b[i] = negmod(addmod(mulmod(a[i], sk_qp[i], q), er[i], q), q);
            }
        }
        for (size_t k = g * alpha; k < (g + 1) * alpha; k++) {
            const uint64_t q = c->q[k];
            uint64_t P = 1;
            for (size_t jj = 0; jj < c->nP; jj++) P = mulmod(P, c->q[c->nQ + jj] % q, q);
            for (size_t j = 0; j < N; j++) {
                size_t i = k * N + j;
                b[i] = addmod(b[i], mulmod(newkey_qp[i], P, q), q);
            }
        }
    }
}

/* secretkey.cu:539-565 gen_relinkey: newkey = s^2. */
void ref_gen_relin_key(const RefCtx *c, const uint64_t *sk_qp, uint64_t seed, uint64_t *key) {
    const size_t N = c->N, QP = c->nQP();
    std::vector<uint64_t> s2(QP * N);
    for (size_t k = 0; k < QP; k++)
        for (size_t j = 0; j < N; j++) s2[k * N + j] = mulmod(sk_qp[k * N + j], sk_qp[k * N + j], c->q[k]);
    ref_gen_kswitch_key(c, sk_qp, s2.data(), seed, key);
}

/* secretkey.cu:577-610 create_galois_keys: newkey = sigma_elt(s) (NTT-domain
 * permutation over QP). */
void ref_gen_galois_key(const RefCtx *c, const uint64_t *sk_qp, uint32_t elt, uint64_t seed, uint64_t *key) {
    const size_t N = c->N, QP = c->nQP();
    std::vector<uint64_t> rs(QP * N);
    ref_apply_galois_ntt(c, elt, sk_qp, rs.data(), QP);
    ref_gen_kswitch_key(c, sk_qp, rs.data(), seed, key);
}

/* secretkey.cu:260-326 encrypt_zero_symmetric (NTT form) + add pt:
 * a from splitmix(seed,0), e from splitmix(seed,1); c0 = -(a s + e) + m, c1 = a. */
void ref_encrypt_symmetric(const RefCtx *c, const uint64_t *sk_qp, const uint64_t *pt, size_t Ql, uint64_t seed,
                           uint64_t *ct) {
    const size_t N = c->N, S = Ql * N;
    auto idx = idx_Ql(*c, Ql, false);
    uint64_t *c0 = ct, *c1 = ct + S;
    ref_sample_uniform(c, sm64(seed, 0), Ql, idx.data(), c1);
    std::vector<int64_t> e(N);
    ref_sample_gaussian(c, sm64(seed, 1), e.data());
    std::vector<uint64_t> er(S);
    for (size_t k = 0; k < Ql; k++)
        for (size_t j = 0; j < N; j++) er[k * N + j] = from_signed(e[j], c->q[k]);
    ntt_limbs(*c, er.data(), Ql, idx.data());
    for (size_t k = 0; k < Ql; k++) {
        const uint64_t q = c->q[k];
        for (size_t j = 0; j < N; j++) {
            size_t i = k * N + j;
            uint64_t v = negmod(addmod(mulmod(c1[i], sk_qp[i], q), er[i], q), q);
            c0[i] = addmod(v, pt[i], q);
        }
    }
}

/* Uniform ternary polynomial (the role of sample_ternary_poly, prng.cu): coefficient
 * j = (splitmix(seed, j) mod 3) - 1. */
void ref_sample_ternary_uniform(const RefCtx *c, uint64_t seed, int64_t *out) {
    for (size_t j = 0; j < c->N; j++) out[j] = (int64_t)(sm64(seed, j) % 3) - 1;
}

/* secretkey.cu:524-535 gen_publickey = encrypt_zero_symmetric (secretkey.cu:260-326)
 * at chain index 0 (all of QP), NTT form: pk = (-(a s + e), a) with a from
 * splitmix(seed, 0) and e from splitmix(seed, 1). pk: [2][QP][N]. */
void ref_gen_public_key(const RefCtx *c, const uint64_t *sk_qp, uint64_t seed, uint64_t *pk) {
    const size_t N = c->N, QP = c->nQP(), S = QP * N;
    auto idx = idx_Ql(*c, c->nQ, true);
    uint64_t *p0 = pk, *p1 = pk + S;
    ref_sample_uniform(c, sm64(seed, 0), QP, idx.data(), p1);
    std::vector<int64_t> e(N);
    ref_sample_gaussian(c, sm64(seed, 1), e.data());
    std::vector<uint64_t> er(S);
    for (size_t k = 0; k < QP; k++)
        for (size_t j = 0; j < N; j++) er[k * N + j] = from_signed(e[j], c->q[k]);
    ntt_limbs(*c, er.data(), QP, idx.data());
    for (size_t k = 0; k < QP; k++) {
        const uint64_t q = c->q[k];
        for (size_t j = 0; j < N; j++) {
            const size_t i = k * N + j;
            p0[i] = negmod(addmod(mulmod(p1[i], sk_qp[i], q), er[i], q), q);
        }
    }
}

/* secretkey.cu:145-217 encrypt_asymmetric (CKKS) with 87-143
 * encrypt_zero_asymmetric_internal and 15-85 ..._internal_internal, for a plaintext
 * at chain index 1: the zero encryption is formed at the key level QP,
 *   c'_i = u pk_i + e   (NTT form, u uniform ternary, i = 0, 1),
 * with ONE error polynomial for both components (the reference draws e for c'_0 and
 * c'_1 from the same seed bytes, secretkey.cu:48-68), then each c'_i is brought to Q
 * by moddown (rns_bconv.cu:712-761: iNTT of the P limbs, fast conversion P -> Q, NTT,
 * (c' - delta) P^-1), and c_0 += m. u from splitmix(seed, 0), e from splitmix(seed, 1).
 * ct: [2][nQ][N]. */
void ref_encrypt_asymmetric(const RefCtx *c, const uint64_t *pk, const uint64_t *pt, uint64_t seed, uint64_t *ct) {
    const size_t N = c->N, QP = c->nQP(), nQ = c->nQ, S = QP * N;
    auto idx = idx_Ql(*c, nQ, true);
    std::vector<int64_t> u(N), e(N);
    ref_sample_ternary_uniform(c, sm64(seed, 0), u.data());
    ref_sample_gaussian(c, sm64(seed, 1), e.data());
    std::vector<uint64_t> ur(S), er(S), cx(S);
    for (size_t k = 0; k < QP; k++)
        for (size_t j = 0; j < N; j++) {
            ur[k * N + j] = from_signed(u[j], c->q[k]);
            er[k * N + j] = from_signed(e[j], c->q[k]);
        }
    ntt_limbs(*c, ur.data(), QP, idx.data());
    ntt_limbs(*c, er.data(), QP, idx.data());
    for (size_t p = 0; p < 2; p++) {
        const uint64_t *pkp = pk + p * S;
        for (size_t k = 0; k < QP; k++) {
            const uint64_t q = c->q[k];
            for (size_t j = 0; j < N; j++) {
                const size_t i = k * N + j;
                cx[i] = addmod(mulmod(ur[i], pkp[i], q), er[i], q);
            }
        }
        ref_moddown(c, cx.data(), nQ, ct + p * nQ * N);
    }
    for (size_t k = 0; k < nQ; k++)
        for (size_t j = 0; j < N; j++) ct[k * N + j] = addmod(ct[k * N + j], pt[k * N + j], c->q[k]);
}

/* secretkey.cu:725-766 ckks_decrypt: m = c0 + c1 s (+ c2 s^2). */
void ref_decrypt(const RefCtx *c, const uint64_t *sk_qp, const uint64_t *ct, size_t size, size_t Ql, uint64_t *pt) {
    const size_t N = c->N, S = Ql * N;
    for (size_t k = 0; k < Ql; k++) {
        const uint64_t q = c->q[k];
        for (size_t j = 0; j < N; j++) {
            size_t i = k * N + j;
            uint64_t s = sk_qp[i];
            uint64_t v = addmod(ct[i], mulmod(ct[S + i], s, q), q);
            if (size > 2) v = addmod(v, mulmod(ct[2 * S + i], mulmod(s, s, q), q), q);
            pt[i] = v;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* CKKS encoder: ckks.cu:89-170 encode_internal, fft.cu:358-439
 * special_fft_backward, rns_base.cu:51-130 decompose_array. Complex products
 * follow cuCmul's operation order without fused multiply-add. */
static inline std::complex<double> cmul(std::complex<double> a, std::complex<double> b) {
    volatile double t0 = a.real() * b.real();
    volatile double t1 = a.imag() * b.imag();
    volatile double t2 = a.real() * b.imag();
    volatile double t3 = a.imag() * b.real();
    return std::complex<double>(t0 - t1, t2 + t3);
}

void ref_encode(const RefCtx *c, const double *re, const double *im, size_t n_values, size_t sparse_slots,
                size_t Ql, double scale, uint64_t *pt) {
    const size_t N = c->N, n = sparse_slots, M = 2 * N;
    int logn = 0;
    while ((size_t(1) << logn) < n) logn++;
    int logM = 0;
    while ((size_t(1) << logM) < M) logM++;
    const int logRatio = logM - logn - 2;
    std::vector<std::complex<double>> in(n, std::complex<double>(0.0, 0.0));
    for (size_t i = 0; i < n; i++) {
        std::complex<double> v = i < n_values ? std::complex<double>(re[i], im ? im[i] : 0.0)
                                              : std::complex<double>(0.0, 0.0);
        in[brv((uint32_t)i, logn)] = v;
    }
    const double fix = scale / (double)n;
    /* GS stages: pairs 1, 2, ..., n/2 (fft.cu:226-296 / 312-356). */
    for (int iter = logn - 1; iter >= 0; iter--) {
        const int logPairs = logn - iter - 1;
        const size_t pairs = size_t(1) << logPairs;
        const bool last = (pairs == n / 2);
        for (size_t tid = 0; tid < n / 2; tid++) {
            size_t k = tid >> logPairs, j = tid & (pairs - 1);
            size_t g = 2 * k * pairs + j;
            uint32_t bk = logn > 1 ? brv((uint32_t)(k << logPairs), logn - 1) : 0;
            uint32_t psiIdx = c->rot_group[bk];
            psiIdx = (uint32_t)(((uint64_t)psiIdx << (logPairs + logRatio)) & (M - 1));
            std::complex<double> tw = c->roots[M - psiIdx];
            std::complex<double> a = in[g], b = in[g + pairs];
            std::complex<double> s0(a.real() + b.real(), a.imag() + b.imag());
            std::complex<double> s1(a.real() - b.real(), a.imag() - b.imag());
            s1 = cmul(s1, tw);
            if (last && fix != 0.0) {
                s0 = std::complex<double>(s0.real() * fix, s0.imag() * fix);
                s1 = std::complex<double>(s1.real() * fix, s1.imag() * fix);
            }
            in[g] = s0;
            in[g + pairs] = s1;
        }
    }
    /* decompose: coefficient ci of the 2n-degree sparse poly sits at
     * ci * (N / 2n); round half away from zero; negative -> q - (|v| mod q). */
    const size_t ratio = N / (2 * n);
    for (size_t k = 0; k < Ql; k++) {
        const uint64_t q = c->q[k];
        uint64_t *dst = pt + k * N;
        std::memset(dst, 0, N * sizeof(uint64_t));
        for (size_t ci = 0; ci < 2 * n; ci++) {
            double v = ci < n ? std::round(in[ci].real()) : std::round(in[ci - n].imag());
            bool neg = std::signbit(v);
            double a = std::fabs(v);
            uint64_t r;
            if (a < 18446744073709551616.0) {
                r = (uint64_t)a % q;
            } else {
                uint64_t lo = (uint64_t)std::fmod(a, 18446744073709551616.0);
                uint64_t hi = (uint64_t)(a / 18446744073709551616.0);
                r = (uint64_t)((((u128)hi << 64) | lo) % q);
            }
            if (neg) r = q - r;
            dst[ci * ratio] = r;
        }
    }
    auto idx = idx_Ql(*c, Ql, false);
    ntt_limbs(*c, pt, Ql, idx.data());
}

/* Constant encode: the special FFT of a constant vector is exact (every
 * difference is zero), so coefficient 0 is round(fl(value * scale)) and all
 * other coefficients are +-0; in NTT form every slot holds [w]_q. */
void ref_encode_scalar(const RefCtx *c, double value, size_t Ql, double scale, uint64_t *pt) {
    std::vector<double> re(c->N / 2, value);
    ref_encode(c, re.data(), nullptr, re.size(), c->N / 2, Ql, scale, pt);
}

/* ckks.cu:172-246 decode_internal + rns_base.cu compose_array + fft.cu:297-356
 * special_fft_forward. */
void ref_decode(const RefCtx *c, const uint64_t *pt, size_t Ql, size_t sparse_slots, double scale, double *re,
                double *im) {
    const size_t N = c->N, n = sparse_slots, M = 2 * N;
    int logn = 0;
    while ((size_t(1) << logn) < n) logn++;
    int logM = 0;
    while ((size_t(1) << logM) < M) logM++;
    const int logRatio = logM - logn - 2;
    std::vector<uint64_t> x(pt, pt + Ql * N);
    auto idx = idx_Ql(*c, Ql, false);
    intt_limbs(*c, x.data(), Ql, idx.data());
    /* multi-precision Q, punctured products, (Q+1)/2 */
    const size_t W = Ql;
    auto mp_mul_small = [&](std::vector<uint64_t> &a, uint64_t b) {
        u128 carry = 0;
        for (size_t i = 0; i < a.size(); i++) {
            u128 t = (u128)a[i] * b + carry;
            a[i] = (uint64_t)t;
            carry = t >> 64;
        }
    };
    auto mp_ge = [&](const uint64_t *a, const uint64_t *b) {
        for (size_t i = W; i-- > 0;) {
            if (a[i] != b[i]) return a[i] > b[i];
        }
        return true;
    };
    std::vector<uint64_t> Q(W, 0);
    Q[0] = 1;
    for (size_t k = 0; k < Ql; k++) mp_mul_small(Q, c->q[k]);
    std::vector<std::vector<uint64_t>> qhat(Ql, std::vector<uint64_t>(W, 0));
    std::vector<uint64_t> qhinv(Ql);
    for (size_t i = 0; i < Ql; i++) {
        qhat[i][0] = 1;
        uint64_t r = 1;
        for (size_t k = 0; k < Ql; k++)
            if (k != i) {
                mp_mul_small(qhat[i], c->q[k]);
                r = mulmod(r, c->q[k] % c->q[i], c->q[i]);
            }
        qhinv[i] = invmod(r, c->q[i]);
    }
    std::vector<uint64_t> half(W);
    {
        /* (Q + 1) / 2 */
        std::vector<uint64_t> t = Q;
        u128 carry = 1;
        for (size_t i = 0; i < W; i++) {
            u128 s = (u128)t[i] + carry;
            t[i] = (uint64_t)s;
            carry = s >> 64;
        }
        for (size_t i = 0; i < W; i++) half[i] = (t[i] >> 1) | (i + 1 < W ? (t[i + 1] << 63) : 0);
    }
    const double inv_scale = 1.0 / scale;
    const double two64 = 18446744073709551616.0;
    const size_t ratio = N / (2 * n);
    std::vector<std::complex<double>> vals(n);
    std::vector<uint64_t> acc(W), tmp(W + 1);
    for (size_t ci = 0; ci < 2 * n; ci++) {
        std::fill(acc.begin(), acc.end(), 0);
        for (size_t i = 0; i < Ql; i++) {
            uint64_t prod = mulmod(x[i * N + ci * ratio], qhinv[i], c->q[i]);
            std::vector<uint64_t> t = qhat[i];
            mp_mul_small(t, prod);
            /* acc = (acc + t) mod Q */
            u128 carry = 0;
            for (size_t w = 0; w < W; w++) {
                u128 s = (u128)acc[w] + t[w] + carry;
                acc[w] = (uint64_t)s;
                carry = s >> 64;
            }
            if (carry || mp_ge(acc.data(), Q.data())) {
                u128 borrow = 0;
                for (size_t w = 0; w < W; w++) {
                    u128 d = (u128)acc[w] - Q[w] - borrow;
                    acc[w] = (uint64_t)d;
                    borrow = (d >> 64) ? 1 : 0;
                }
            }
        }
        double res = 0.0, sc = inv_scale;
        if (mp_ge(acc.data(), half.data())) {
            for (size_t w = 0; w < W; w++, sc *= two64) {
                if (acc[w] > Q[w]) {
                    uint64_t d = acc[w] - Q[w];
                    res += d ? (double)d * sc : 0.0;
                } else {
                    uint64_t d = Q[w] - acc[w];
                    res -= d ? (double)d * sc : 0.0;
                }
            }
        } else {
            for (size_t w = 0; w < W; w++, sc *= two64) {
                uint64_t d = acc[w];
                res += d ? (double)d * sc : 0.0;
            }
        }
        if (ci < n)
            vals[ci] = std::complex<double>(res, vals[ci].imag());
        else
            vals[ci - n] = std::complex<double>(vals[ci - n].real(), res);
    }
    /* CT stages: pairs n/2, ..., 1 (fft.cu:88-224). */
    for (int iter = 0; iter < logn; iter++) {
        const int logPairs = logn - iter - 1;
        const size_t pairs = size_t(1) << logPairs;
        for (size_t tid = 0; tid < n / 2; tid++) {
            size_t k = tid >> logPairs, j = tid & (pairs - 1);
            size_t g = 2 * k * pairs + j;
            uint32_t bk = logn > 1 ? brv((uint32_t)(k << logPairs), logn - 1) : 0;
            uint32_t psiIdx = c->rot_group[bk];
            psiIdx = (uint32_t)(((uint64_t)psiIdx << (logPairs + logRatio)) & (M - 1));
            std::complex<double> tw = c->roots[psiIdx];
            std::complex<double> s0 = vals[g], s1 = cmul(vals[g + pairs], tw);
            vals[g] = std::complex<double>(s0.real() + s1.real(), s0.imag() + s1.imag());
            vals[g + pairs] = std::complex<double>(s0.real() - s1.real(), s0.imag() - s1.imag());
        }
    }
    for (size_t i = 0; i < n; i++) {
        auto v = vals[i];
        size_t r = brv((uint32_t)i, logn);
        re[r] = v.real();
        if (im) im[r] = v.imag();
    }
}

/* Ct_pt_matrix_mul.cuh:18-185 ct_pt_matrix_mul_wo_pre: out_i = rescale(
 * sum_j X_j (.) encode(W[j][i], scale_j)); each constant plaintext is
 * [round(fl(W[j][i] * scale_j))]_q in every NTT slot. X_j: 2 x Ql x N;
 * W row-major row_w x col_w; out[i - col_begin]: 2 x (Ql-1) x N. */
void ref_ctpt_matmul(const RefCtx *c, const uint64_t *const *X, size_t row_w, const double *W, size_t col_w,
                     size_t Ql, const double *scales, size_t col_begin, size_t col_end, uint64_t *const *out) {
    const size_t N = c->N, S = Ql * N;
    for (size_t i = col_begin; i < col_end; i++) {
        std::vector<uint64_t> acc(2 * S, 0);
        std::vector<uint64_t> w(row_w * Ql), ws(row_w * Ql);
        for (size_t j = 0; j < row_w; j++) {
            double v = std::round(W[j * col_w + i] * scales[j]);
            bool neg = std::signbit(v);
            double a = std::fabs(v);
            for (size_t k = 0; k < Ql; k++) {
                const uint64_t q = c->q[k];
                uint64_t r;
                if (a < 18446744073709551616.0) {
                    r = (uint64_t)a % q;
                } else {
                    uint64_t lo = (uint64_t)std::fmod(a, 18446744073709551616.0);
                    uint64_t hi = (uint64_t)(a / 18446744073709551616.0);
                    r = (uint64_t)((((u128)hi << 64) | lo) % q);
                }
                if (neg && r) r = q - r;
                w[j * Ql + k] = r;
                ws[j * Ql + k] = shoup(r, q);
            }
        }
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (long pk = 0; pk < (long)(2 * Ql); pk++) {
            const size_t k = (size_t)pk % Ql;
            const uint64_t q = c->q[k];
            uint64_t *a = acc.data() + pk * N;
            for (size_t j = 0; j < row_w; j++) {
                const uint64_t *x = X[j] + pk * N;
                const uint64_t wv = w[j * Ql + k], wsv = ws[j * Ql + k];
                for (size_t n = 0; n < N; n++) a[n] = addmod(a[n], mul_shoup(x[n], wv, wsv, q), q);
            }
        }
        ref_rescale(c, acc.data(), out[i - col_begin], Ql, 2);
    }
}


/* Bootstrapper.cu:3271-3284 kernel_modraise_inplace, 3287-3356 modraise_inplace:
 * in = [2][N] (prime 0, NTT form) -> out = [2][nQ][N] (NTT form): iNTT, lift each
 * coefficient x of [0, q0) to x (x <= q0/2) or x - q0 (x > q0/2) mod every data
 * prime, NTT. */
void ref_modraise(const RefCtx *c, const uint64_t *in, uint64_t *out) {
    const size_t N = c->N, nq = c->nQ;
    const uint64_t q0 = c->q[0];
    for (size_t p = 0; p < 2; p++) {
        std::vector<uint64_t> a(in + p * N, in + (p + 1) * N);
        intt_limb(*c, a.data(), 0);
        for (size_t k = 0; k < nq; k++) {
            const uint64_t q = c->q[k];
            uint64_t *o = out + (p * nq + k) * N;
            for (size_t i = 0; i < N; i++) {
                uint64_t v = a[i] % q;
                if (a[i] > (q0 >> 1)) v = submod(v, q0 % q, q);
                o[i] = v;
            }
            ntt_limb(*c, o, k);
        }
    }
}

/* Multiply npolys x [nlimbs][N] NTT-form polynomials (primes 0..nlimbs-1) by the
 * monomial X^power (0 <= power < 2N) in the coefficient domain: negacyclic shift. */
void ref_mul_monomial(const RefCtx *c, uint64_t *data, size_t nlimbs, size_t npolys, size_t power) {
    const size_t N = c->N;
    for (size_t p = 0; p < npolys; p++) {
        for (size_t k = 0; k < nlimbs; k++) {
            uint64_t *a = data + (p * nlimbs + k) * N;
            const uint64_t q = c->q[k];
            intt_limb(*c, a, k);
            std::vector<uint64_t> r(N);
            for (size_t i = 0; i < N; i++) {
                const size_t e = (i + power) % (2 * N);
                const uint64_t v = a[i];
                if (e < N)
                    r[e] = v;
                else
                    r[e - N] = v ? q - v : 0;
            }
            std::copy(r.begin(), r.end(), a);
            ntt_limb(*c, a, k);
        }
    }
}


/* ---- shared-a format (FmtSwitch.cu, ShRotate.cu) --------------------------------------- */

namespace {
uint64_t prod_special_mod(const RefCtx *c, uint64_t q) {
    uint64_t P = 1;
    for (size_t jj = 0; jj < c->nP; jj++) P = mulmod(P, c->q[c->nQ + jj] % q, q);
    return P;
}
}  // namespace

/* FmtSwitch.cu:82-108 extend_Ql_to_P_kernel: coefficient residues x_g (g < Ql) ->
 * residues mod every special prime of x = sum_g y_g QHat_g - v Ql, y_g =
 * [x_g QHat_g^-1]_{q_g}, v = floor(sum_g y_g / q_g) summed in double in g order,
 * rounded up when within 1e-9 below an integer. outP: [nP][N]. */
void ref_extend_exact(const RefCtx *c, const uint64_t *coef, size_t Ql, uint64_t *outP) {
    const size_t N = c->N, nP = c->nP;
    std::vector<uint64_t> inv(Ql);
    std::vector<uint32_t> S;
    for (size_t g = 0; g < Ql; g++) S.push_back((uint32_t)g);
    for (size_t g = 0; g < Ql; g++) inv[g] = invmod(punctured_prod_mod(*c, S, g, c->q[g]), c->q[g]);
    for (size_t j = 0; j < N; j++) {
        std::vector<uint64_t> y(Ql);
        double v = 0.0;
        for (size_t g = 0; g < Ql; g++) {
            y[g] = mulmod(coef[g * N + j], inv[g], c->q[g]);
            v += (double)y[g] / (double)c->q[g];
        }
        uint64_t alpha = (uint64_t)v;
        if (v - (double)alpha > 1.0 - 1e-9) alpha += 1;
        for (size_t k = 0; k < nP; k++) {
            const uint64_t p = c->q[c->nQ + k];
            u128 acc = 0;
            for (size_t g = 0; g < Ql; g++) acc += (u128)y[g] * punctured_prod_mod(*c, S, g, p);
            uint64_t Qlp = 1;
            for (size_t g = 0; g < Ql; g++) Qlp = mulmod(Qlp, c->q[g] % p, p);
            outP[k * N + j] = submod((uint64_t)(acc % p), mulmod(alpha % p, Qlp, p), p);
        }
    }
}

/* FmtSwitch.cu:164-270 gen_format_key with injected randomness (fseed =
 * splitmix(seed, n); a_fmt[j] uniform from splitmix(fseed, j); e_{j,i} from
 * splitmix(fseed, n + j n + i)): over QlP (local limb order),
 * B[j][i] = -(a_fmt[j] s'_i + e_{j,i}) + [i == j] P s on the Ql limbs. */
void ref_gen_fmt_key(const RefCtx *c, const uint64_t *sk_qp, const uint64_t *const *aux_qp, size_t n, size_t Ql,
                     uint64_t seed, uint64_t *a_fmt, uint64_t *B) {
    const size_t N = c->N, QlP = Ql + c->nP;
    auto idx = idx_Ql(*c, Ql, true);
    const uint64_t fseed = sm64(seed, n);
    std::vector<int64_t> e(N);
    std::vector<uint64_t> er(QlP * N);
    for (size_t j = 0; j < n; j++) {
        uint64_t *aj = a_fmt + j * QlP * N;
        ref_sample_uniform(c, sm64(fseed, j), QlP, idx.data(), aj);
        for (size_t i = 0; i < n; i++) {
            ref_sample_gaussian(c, sm64(fseed, n + j * n + i), e.data());
            for (size_t k = 0; k < QlP; k++)
                for (size_t t = 0; t < N; t++) er[k * N + t] = from_signed(e[t], c->q[idx[k]]);
            ntt_limbs(*c, er.data(), QlP, idx.data());
            uint64_t *b = B + (j * n + i) * QlP * N;
            for (size_t k = 0; k < QlP; k++) {
                const uint32_t pi = idx[k];
                const uint64_t q = c->q[pi], Pq = prod_special_mod(c, q);
                for (size_t t = 0; t < N; t++) {
                    uint64_t v = negmod(addmod(mulmod(aj[k * N + t], aux_qp[i][pi * N + t], q), er[k * N + t], q), q);
                    if (i == j && k < Ql) v = addmod(v, mulmod(sk_qp[pi * N + t], Pq, q), q);
                    b[k * N + t] = v;
                }
            }
        }
    }
}

/* FmtSwitch.cu:276-394 fmt_switch_forward: cts n x [2][Ql][N] -> out [n+1][Ql][N]
 * (b'_0..b'_{n-1}, a'). */
void ref_fmt_switch_forward(const RefCtx *c, const uint64_t *const *cts, size_t n, size_t Ql, const uint64_t *a_fmt,
                            const uint64_t *B, uint64_t *out) {
    const size_t N = c->N, nP = c->nP, QlP = Ql + nP;
    auto idx = idx_Ql(*c, Ql, true);
    auto idxQ = idx_Ql(*c, Ql, false);
    std::vector<uint32_t> pP;
    for (size_t k = 0; k < nP; k++) pP.push_back((uint32_t)(c->nQ + k));
    std::vector<uint64_t> ext(n * QlP * N);
    for (size_t j = 0; j < n; j++) {
        const uint64_t *aj = cts[j] + Ql * N;
        uint64_t *ej = ext.data() + j * QlP * N;
        std::memcpy(ej, aj, Ql * N * sizeof(uint64_t));
        std::vector<uint64_t> coef(aj, aj + Ql * N);
        intt_limbs(*c, coef.data(), Ql, idxQ.data());
        ref_extend_exact(c, coef.data(), Ql, ej + Ql * N);
        ntt_limbs(*c, ej + Ql * N, nP, pP.data());
    }
    std::vector<uint64_t> cx(QlP * N), md(Ql * N);
    for (size_t o = 0; o <= n; o++) {
        for (size_t k = 0; k < QlP; k++) {
            const uint64_t q = c->q[idx[k]];
            for (size_t t = 0; t < N; t++) {
                u128 acc = 0;
                for (size_t j = 0; j < n; j++) {
                    const uint64_t *key = o < n ? B + (j * n + o) * QlP * N : a_fmt + j * QlP * N;
                    acc += (u128)ext[(j * QlP + k) * N + t] * key[k * N + t];
                }
                cx[k * N + t] = (uint64_t)(acc % q);
            }
        }
        ref_moddown(c, cx.data(), Ql, md.data());
        uint64_t *dst = out + o * Ql * N;
        for (size_t k = 0; k < Ql; k++)
            for (size_t t = 0; t < N; t++) {
                const size_t i = k * N + t;
                dst[i] = o < n ? addmod(cts[o][i], md[i], c->q[k]) : md[i];
            }
    }
}

/* ShRotate.cu:54-162 gen_shared_galois_key with injected randomness (kseed =
 * splitmix(seed, elt); A_g from splitmix(kseed, g (n+1)); e_{g,i} from
 * splitmix(kseed, g (n+1) + 1 + i)): A [dnum][QP][N], B [dnum][n][QP][N] with
 * B_{g,i} = -(A_g s'_i + e) + P [sigma(s'_i)]_{digit g}. */
void ref_gen_shared_galois_key(const RefCtx *c, const uint64_t *const *aux_qp, size_t n, uint32_t elt, uint64_t seed,
                               uint64_t *A, uint64_t *B) {
    const size_t N = c->N, QP = c->nQP(), alpha = c->nP, dnum = c->nQ / c->nP;
    auto idx = idx_Ql(*c, c->nQ, true);
    const uint64_t kseed = sm64(seed, elt);
    std::vector<std::vector<uint64_t>> sig(n, std::vector<uint64_t>(QP * N));
    for (size_t i = 0; i < n; i++) ref_apply_galois_ntt(c, elt, aux_qp[i], sig[i].data(), QP);
    std::vector<int64_t> e(N);
    std::vector<uint64_t> er(QP * N);
    for (size_t g = 0; g < dnum; g++) {
        uint64_t *Ag = A + g * QP * N;
        ref_sample_uniform(c, sm64(kseed, g * (n + 1)), QP, idx.data(), Ag);
        for (size_t i = 0; i < n; i++) {
            ref_sample_gaussian(c, sm64(kseed, g * (n + 1) + 1 + i), e.data());
            for (size_t k = 0; k < QP; k++)
                for (size_t t = 0; t < N; t++) er[k * N + t] = from_signed(e[t], c->q[k]);
            ntt_limbs(*c, er.data(), QP, idx.data());
            uint64_t *b = B + (g * n + i) * QP * N;
            for (size_t k = 0; k < QP; k++) {
                const uint64_t q = c->q[k];
                for (size_t t = 0; t < N; t++) {
                    const size_t x = k * N + t;
                    b[x] = negmod(addmod(mulmod(Ag[x], aux_qp[i][x], q), er[x], q), q);
                }
            }
            for (size_t k = g * alpha; k < (g + 1) * alpha; k++) {
                const uint64_t q = c->q[k], Pq = prod_special_mod(c, q);
                for (size_t t = 0; t < N; t++) {
                    const size_t x = k * N + t;
                    b[x] = addmod(b[x], mulmod(sig[i][x], Pq, q), q);
                }
            }
        }
    }
}

}  // extern "C"
