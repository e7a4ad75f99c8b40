// Exact 64-bit modular arithmetic for CDNA4 (gfx950) kernels.
//
// CDNA4 has no 64x64->128 multiply instruction; __umul64hi and 64-bit
// products lower to 32-bit v_mul_hi/v_mad_u64_u32 chains. Every helper keeps
// operands below 2^62 so that lazy [0, 4q) butterflies never wrap (all primes
// are at most 61 bits).
//
// The ring degree N is a power of two (2^10 .. 2^17), so kernels split a flat
// index over [..][N] with a shift and a mask rather than a 64-bit division.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace shm {

typedef unsigned __int128 u128;

// One RNS prime with its Barrett constant floor(2^128 / q) = (mu_hi, mu_lo).
struct DevMod {
    uint64_t q;
    uint64_t mu_hi;
    uint64_t mu_lo;
};

__device__ __forceinline__ uint64_t csub(uint64_t x, uint64_t q) { return x >= q ? x - q : x; }

__device__ __forceinline__ uint64_t add_mod(uint64_t a, uint64_t b, uint64_t q) { return csub(a + b, q); }

__device__ __forceinline__ uint64_t sub_mod(uint64_t a, uint64_t b, uint64_t q) { return csub(a + q - b, q); }

__device__ __forceinline__ uint64_t neg_mod(uint64_t a, uint64_t q) { return a == 0 ? 0 : q - a; }

// Shoup product x * w mod q with ws = floor(w * 2^64 / q): result in [0, 2q)
// for any x < 2^64.
__device__ __forceinline__ uint64_t mul_shoup_lazy(uint64_t x, uint64_t w, uint64_t ws, uint64_t q) {
    return x * w - __umul64hi(x, ws) * q;
}

// x w mod q in [0, 4q) for any x < 2^64: the quotient estimate keeps x_hi ws_hi and the
// high halves of the two middle partial products and drops the lowest one, so it is at
// most 2 below Shoup's floor(x ws / 2^64) (two 32-bit high multiplies and one
// 32 x 32 + 64 multiply-add instead of a full 64 x 64 high product).
__device__ __forceinline__ uint64_t mul_shoup_lazy4(uint64_t x, uint64_t w, uint64_t ws, uint64_t q) {
    const uint32_t xl = (uint32_t)x, xh = (uint32_t)(x >> 32);
    const uint32_t sl = (uint32_t)ws, sh = (uint32_t)(ws >> 32);
    const uint64_t mid = (uint64_t)__umulhi(xl, sh) + (uint64_t)__umulhi(xh, sl);
    const uint64_t qe = (uint64_t)xh * sh + mid;
    return x * w - qe * q;
}

__device__ __forceinline__ uint64_t mul_shoup(uint64_t x, uint64_t w, uint64_t ws, uint64_t q) {
    return csub(mul_shoup_lazy(x, w, ws, q), q);
}

// Shoup companion floor(w * 2^64 / q) of w < q from the Barrett constant:
// floor(w * mu / 2^64) is at most one short, and the remainder w * 2^64 - e * q
// (exact in 64 bits, below 2q) says when.
__device__ __forceinline__ uint64_t shoup_companion(uint64_t w, const DevMod &m) {
    uint64_t e = w * m.mu_hi + __umul64hi(w, m.mu_lo);
    const uint64_t rem = 0 - e * m.q;
    return rem >= m.q ? e + 1 : e;
}

// Barrett reduction of a 128-bit value (hi, lo) with mu = floor(2^128/q).
// The quotient estimate is at most 2 short, so the remainder is in [0, 3q)
// before the final corrections.
__device__ __forceinline__ uint64_t barrett128(uint64_t hi, uint64_t lo, const DevMod &m) {
    // floor((hi*2^64 + lo) * mu / 2^128), dropping the lo*mu_lo low part.
    const uint64_t a = __umul64hi(lo, m.mu_lo);
    const uint64_t b_lo = lo * m.mu_hi, b_hi = __umul64hi(lo, m.mu_hi);
    const uint64_t c_lo = hi * m.mu_lo, c_hi = __umul64hi(hi, m.mu_lo);
    uint64_t s = a + b_lo;
    uint64_t carry = s < a;
    const uint64_t s2 = s + c_lo;
    carry += s2 < s;
    const uint64_t qest = hi * m.mu_hi + b_hi + c_hi + carry;
    uint64_t r = lo - qest * m.q;
    r = csub(r, m.q);
    return csub(r, m.q);
}

__device__ __forceinline__ uint64_t mul_mod(uint64_t a, uint64_t b, const DevMod &m) {
    return barrett128(__umul64hi(a, b), a * b, m);
}

// Reduce a single 64-bit word (any value) mod q using mu_hi = floor(2^64/q).
__device__ __forceinline__ uint64_t reduce64(uint64_t x, const DevMod &m) {
    const uint64_t qe = __umul64hi(x, m.mu_hi);
    return csub(x - qe * m.q, m.q);
}

// 128-bit accumulator helpers (explicit carries; no __int128 division).
struct Acc128 {
    uint64_t lo, hi;
};

__device__ __forceinline__ void mac128(Acc128 &acc, uint64_t a, uint64_t b) {
    const uint64_t lo = a * b, hi = __umul64hi(a, b);
    acc.lo += lo;
    acc.hi += hi + (acc.lo < lo);
}

__device__ __forceinline__ uint64_t reduce128(const Acc128 &acc, const DevMod &m) { return barrett128(acc.hi, acc.lo, m); }

// Sums of products of residues below 2^52 (every prime in (2^45, 2^52)): with
// a = a1 2^26 + a0 and b = b1 2^26 + b0, a b = a1 b1 2^52 + (a1 b0 + a0 b1) 2^26 + a0 b0,
// each partial product a 32 x 32 -> 64-bit multiply-add into one of three 64-bit
// sums (no carries: up to 64 terms stay below 2^64).
struct Acc52 {
    uint64_t s0, s1, s2;
};

struct Split26 {
    uint32_t hi, lo;
};

__device__ __forceinline__ Split26 split26(uint64_t x) { return {(uint32_t)(x >> 26), (uint32_t)x & 0x3ffffffu}; }

__device__ __forceinline__ void mac52(Acc52 &acc, Split26 a, uint64_t b) {
    const Split26 bs = split26(b);
    acc.s0 += (uint64_t)a.lo * bs.lo;
    acc.s1 += (uint64_t)a.hi * bs.lo;
    acc.s1 += (uint64_t)a.lo * bs.hi;
    acc.s2 += (uint64_t)a.hi * bs.hi;
}

// [S]_q of S = s2 2^52 + s1 2^26 + s0 < 2^109 for 2^45 < q < 2^52: T = floor(S / 2^45)
// exactly, the quotient estimate floor(T mu / 2^64) with mu = floor(2^109 / q) (from
// the Barrett constant floor(2^128 / q)) is at most 3 short, so S - qe q in [0, 4q)
// is taken in 64 bits.
__device__ __forceinline__ uint64_t reduce52(const Acc52 &a, const DevMod &m) {
    const uint64_t T = (a.s2 << 7) + (a.s1 >> 19) + ((((a.s1 & 0x7ffffull) << 26) + a.s0) >> 45);
    const uint64_t lo = a.s0 + (a.s1 << 26) + (a.s2 << 52);
    const uint64_t mu = (m.mu_hi << 45) | (m.mu_lo >> 19);
    const uint64_t r = lo - __umul64hi(T, mu) * m.q;
    return csub(csub(r, m.q << 1), m.q);
}

// a b mod q for a, b < 2^53 and 2^45 < q < 2^52: the four 26-bit split products (S < 2^106)
// and reduce52, about half the multiplies of mul_mod; the same canonical residue.
__device__ __forceinline__ uint64_t mul_mod52(uint64_t a, uint64_t b, const DevMod &m) {
    Acc52 acc{0, 0, 0};
    mac52(acc, split26(a), b);
    return reduce52(acc, m);
}
// [v]_q of an integral double v with |v| < 2^128; negative v maps to q - (|v| mod q),
// which is q itself when |v| is a multiple of q (callers reduce when they need [0, q)).
__device__ __forceinline__ uint64_t residue_of_rounded(double v, const DevMod &m) {
    const bool neg = signbit(v);
    const double a = fabs(v);
    uint64_t r;
    if (a < 18446744073709551616.0) {
        r = reduce64((uint64_t)a, m);
    } else {
        const uint64_t lo = (uint64_t)fmod(a, 18446744073709551616.0);
        const uint64_t hi = (uint64_t)(a / 18446744073709551616.0);
        r = barrett128(hi, lo, m);
    }
    return neg ? m.q - r : r;
}

}  // namespace shm
