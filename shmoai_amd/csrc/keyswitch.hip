// Hybrid (dnum) key switching: ModUp -> key inner product -> ModDown.
//
// ModUp (rns_bconv.cu:530-634): c2 (Ql limbs, NTT) -> per digit g of alpha
// limbs: own limbs copied in NTT form; complement limbs of QlP get the
// approximate fast base conversion sum_i [x_i qhat_i^-1]_{q_i} [qhat_i]_p of the
// coefficient-form digit (alpha = 1: [x]_p), then an NTT.
// Inner product (eval_key_switch.cu:12-60): over QlP limbs,
// cx0 = sum_g d_g key_g.b, cx1 = sum_g d_g key_g.a with 128-bit accumulation.
// ModDown (rns_bconv.cu:776-828, ntt_moddown.cu:106-261): iNTT of the P limbs,
// delta = fastbconv(P -> Ql), NTT(delta) whose last pass writes
// out = (cx - delta) P^-1 mod q (adding onto the output where asked).
#include <algorithm>

#include <atomic>
#include <cstdio>
#include <functional>
#include <map>
#include <mutex>
#include <string>

#include <type_traits>

#include "engine.h"

namespace shm {

namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(size_t work) {
    size_t g = (work + kBlock - 1) / kBlock;
    if (g > (1u << 20)) g = 1u << 20;
    return (unsigned)(g ? g : 1);
}

__device__ __forceinline__ uint32_t qlp_prime(size_t k, size_t Ql, size_t nQ) {
    return (uint32_t)(k < Ql ? k : nQ + (k - Ql));
}

// One thread per (digit g, coefficient j): the digit's own limbs are copied (in
// coefficient form, transformed with the rest) and its y_i = [x_i (D_g/q_i)^-1]_{q_i} are formed once, then every other
// QlP limb gets sum_i y_i [D_g/q_i]_p (the approximate fast conversion). With
// `rounded`, the conversion is exact and centred (it subtracts
// round(sum_i y_i / q_i) D_g), so every digit is lifted to (-D_g/2, D_g/2]:
// zero-mean digits keep the key-error product free of the constant component
// that otherwise lands on the slots nearest w = 1.
constexpr int kMaxAlpha = 16;

// The base conversions unroll their sums over the digit's (or the special primes')
// limbs to a compile-time bound A >= that count, so the common sizes (5 special primes
// at the bootstrapping parameters, 1 for the matmul's chain) carry no tests for absent
// terms: A is 1, 2, 4, 5, 8 or 16.
template <typename F>
void with_alpha_bound(size_t alpha, F &&f) {
    if (alpha <= 1)
        f(std::integral_constant<int, 1>{});
    else if (alpha <= 2)
        f(std::integral_constant<int, 2>{});
    else if (alpha <= 4)
        f(std::integral_constant<int, 4>{});
    else if (alpha == 5)
        f(std::integral_constant<int, 5>{});
    else if (alpha <= 8)
        f(std::integral_constant<int, 8>{});
    else
        f(std::integral_constant<int, kMaxAlpha>{});
}

// Blocks never straddle two digits (N is a multiple of the block size), so the
// digit index and every table address derived from it are wave-uniform.
// `members` polynomials are converted at once: member m reads tcoef + m Ql N
// and writes out + m beta QlP N.
template <bool P52, int A>
__global__ void k_modup_bconv(const uint64_t *__restrict__ tcoef_all, uint64_t *__restrict__ out_all,
                              const DevMod *__restrict__ mods, const uint64_t *__restrict__ hat_inv,
                              const uint64_t *__restrict__ hat_inv_s, const uint64_t *__restrict__ hat_mod,
                              const uint64_t *__restrict__ dig_mod, size_t Ql, size_t nQ, size_t alpha, size_t beta,
                              size_t N, size_t members, int rounded, int write_own) {
    const size_t QlP = Ql + alpha;
    const size_t total = members * beta * N;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t gm = __builtin_amdgcn_readfirstlane((uint32_t)(t >> __builtin_ctzll(N))), j = t & (N - 1);
        const size_t mem = gm / beta, g = gm % beta;
        const uint64_t *tcoef = tcoef_all + mem * Ql * N;
        const size_t start = g * alpha, size = min(alpha, Ql - start);
        uint64_t *o = out_all + (mem * beta + g) * QlP * N + j;
        uint64_t y[A];
        double v = 0.0;
#pragma unroll
        for (int i = 0; i < A; i++) {
            y[i] = 0;
            if ((size_t)i < size) {
                const size_t li = start + i;
                const uint64_t qi = mods[li].q;
                y[i] = size == 1 ? tcoef[li * N + j] : mul_shoup(tcoef[li * N + j], hat_inv[li], hat_inv_s[li], qi);
                if (rounded) v = __dadd_rn(v, __ddiv_rn((double)y[i], (double)qi));
                // own limb in coefficient form: the NTT below restores c2 exactly
                // (without write_own the caller reads the own limbs from c2 itself)
                if (write_own) o[li * N] = tcoef[li * N + j];
            }
        }
        const uint64_t r = rounded ? (uint64_t)__dadd_rn(v, 0.5) : 0;
        const uint64_t qs = mods[start].q;
        Split26 ys[A];
#pragma unroll
        for (int i = 0; i < A; i++) ys[i] = split26(P52 ? y[i] : 0);
        for (size_t k = 0; k < QlP; k++) {
            if (k >= start && k < start + size) continue;
            const DevMod mp = mods[qlp_prime(k, Ql, nQ)];
            uint64_t val;
            if (size == 1) {
                val = reduce64(y[0], mp);
                if (r) val = sub_mod(val, reduce64(qs, mp), mp.q);
            } else if (P52) {
                Acc52 acc{0, 0, 0};
                const uint64_t *hm = hat_mod + (g * alpha) * QlP + k;
                if (size == (size_t)A) {
                    // a full digit (every digit but possibly the last): no per-term tests,
                    // so the term constants are fetched together
#pragma unroll
                    for (int i = 0; i < A; i++) mac52(acc, ys[i], hm[i * QlP]);
                } else {
#pragma unroll
                    for (int i = 0; i < A; i++)
                        if ((size_t)i < size) mac52(acc, ys[i], hm[i * QlP]);
                }
                if (rounded) mac52(acc, split26(mp.q - r), dig_mod[g * QlP + k]);
                val = reduce52(acc, mp);
            } else {
                Acc128 acc{0, 0};
                const uint64_t *hm = hat_mod + (g * alpha) * QlP + k;
#pragma unroll
                for (int i = 0; i < A; i++)
                    if ((size_t)i < size) mac128(acc, y[i], hm[i * QlP]);
                // - r D_g == (p - r) [D_g]_p (mod p), folded in before the one reduction
                if (rounded) mac128(acc, mp.q - r, dig_mod[g * QlP + k]);
                val = reduce128(acc, mp);
            }
            __builtin_nontemporal_store(val, &o[k * N]);
        }
    }
}

__global__ void k_ks_inner(const uint64_t *__restrict__ d, const uint64_t *__restrict__ key,
                           uint64_t *__restrict__ cx, const DevMod *__restrict__ mods, size_t Ql, size_t nQ,
                           size_t QP, size_t alpha, size_t beta, size_t N) {
    const size_t QlP = Ql + alpha, total = QlP * N;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t k = t >> __builtin_ctzll(N), j = t & (N - 1);
        const uint32_t pi = qlp_prime(k, Ql, nQ);
        Acc128 a0{0, 0}, a1{0, 0};
        for (size_t g = 0; g < beta; g++) {
            const uint64_t dv = d[g * total + t];
            const uint64_t *kg = key + g * 2 * QP * N;
            mac128(a0, dv, kg[(size_t)pi * N + j]);
            mac128(a1, dv, kg[(QP + pi) * N + j]);
        }
        const DevMod m = mods[pi];
        cx[t] = reduce128(a0, m);
        cx[total + t] = reduce128(a1, m);
    }
}

// cx[o][k][j] = sum_g d[g][k][j] * key_o(g)[limb(k)][j] for every output o. Key
// limbs are addressed by their global prime (QP layout) or, with `local`, by the
// QlP position (keys generated for one level).
// Layout: one thread per (output, limb k, coefficient j); a block covers
// J = 256 / OG consecutive coefficients and OG = 2^lgog output slots (OG <= 16).
// The OG threads of one coefficient load the same digit words (L1 serves all
// but the first), so the digits leave HBM once for all outputs, and every
// thread waits only on its own 2 ndig loads.
template <int MAXB, bool P52>
__global__ __launch_bounds__(256) void k_ks_inner_multi(const uint64_t *__restrict__ d, KeyRefs kr, int nout,
                                                        uint64_t *__restrict__ cx, const DevMod *__restrict__ mods,
                                                        size_t Ql, size_t nQ, size_t alpha, int ndig, size_t N,
                                                        int local, const uint32_t *__restrict__ perm, int accumulate,
                                                        int lgog) {
    const size_t QlP = Ql + alpha, total = QlP * N;
    const int lgj = 8 - lgog;
    const size_t t = ((size_t)blockIdx.x << lgj) + (threadIdx.x & ((1u << lgj) - 1));
    const int os = (int)(threadIdx.x >> lgj);
    if (t >= total || os >= nout) return;
    const size_t k = t >> __builtin_ctzll(N), j = t & (N - 1);
    const uint32_t pi = qlp_prime(k, Ql, nQ);
    const size_t kl = (local ? k : (size_t)pi) * N + j;
    // hoisted rotation: the digits are read through the NTT-domain Galois permutation
    const size_t src = perm ? k * N + perm[j] : t;
    uint64_t dv[MAXB];
#pragma unroll
    for (int g = 0; g < MAXB; g++) dv[g] = g < ndig ? d[(size_t)g * total + src] : 0;
    Split26 ds[MAXB];
#pragma unroll
    for (int g = 0; g < MAXB; g++) ds[g] = split26(P52 ? dv[g] : 0);
    const DevMod m = mods[pi];
    // two outputs per step: their 2 ndig key loads are in flight together
    const int step = 1 << lgog;
    for (int o = os; o < nout; o += 2 * step) {
        const int o2 = o + step;
        const bool two = o2 < nout;
        const uint64_t *kb = kr.base[o], *kb2 = kr.base[two ? o2 : o];
        const size_t st = kr.stride[o], st2 = kr.stride[two ? o2 : o];
        uint64_t kv[MAXB], kv2[MAXB];
#pragma unroll
        for (int g = 0; g < MAXB; g++) {
            kv[g] = g < ndig ? __builtin_nontemporal_load(&kb[(size_t)g * st + kl]) : 0;
            kv2[g] = g < ndig && two ? __builtin_nontemporal_load(&kb2[(size_t)g * st2 + kl]) : 0;
        }
        uint64_t v, v2;
        if (P52) {
            Acc52 a{0, 0, 0}, a2{0, 0, 0};
#pragma unroll
            for (int g = 0; g < MAXB; g++)
                if (g < ndig) {
                    mac52(a, ds[g], kv[g]);
                    mac52(a2, ds[g], kv2[g]);
                }
            v = reduce52(a, m);
            v2 = reduce52(a2, m);
        } else {
            Acc128 a{0, 0}, a2{0, 0};
#pragma unroll
            for (int g = 0; g < MAXB; g++)
                if (g < ndig) {
                    mac128(a, dv[g], kv[g]);
                    mac128(a2, dv[g], kv2[g]);
                }
            v = reduce128(a, m);
            v2 = reduce128(a2, m);
        }
        __builtin_nontemporal_store(accumulate ? add_mod(cx[(size_t)o * total + t], v, m.q) : v, &cx[(size_t)o * total + t]);
        if (two)
            __builtin_nontemporal_store(accumulate ? add_mod(cx[(size_t)o2 * total + t], v2, m.q) : v2,
                                        &cx[(size_t)o2 * total + t]);
    }
}

// delta[p][i][j] = fastbconv of the coefficient-form P limbs of polynomial p to
// q_i, one thread per (p, j). With `rounded`, the conversion is exact and
// centred: it subtracts round(sum_j y_j / p_j) P, so delta is the representative
// of cx mod P in (-P/2, P/2] and (cx - delta) / P rounds to nearest.
template <bool P52, int A>
__global__ void k_moddown_bconv(const uint64_t *__restrict__ cx, uint64_t *__restrict__ delta,
                                const DevMod *__restrict__ mods, const uint64_t *__restrict__ phat_inv,
                                const uint64_t *__restrict__ phat_inv_s, const uint64_t *__restrict__ phat_mod,
                                const uint64_t *__restrict__ pmod, size_t Ql, size_t nQ, size_t nP, size_t N,
                                size_t npolys, int rounded) {
    const size_t QlP = Ql + nP, total = npolys * N;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t p = t >> __builtin_ctzll(N), j = t & (N - 1);
        const uint64_t *cxP = cx + (p * QlP + Ql) * N + j;
        uint64_t *d = delta + p * Ql * N + j;
        uint64_t y[A];
        double v = 0.0;
#pragma unroll
        for (int jj = 0; jj < A; jj++) {
            y[jj] = 0;
            if ((size_t)jj < nP) {
                const uint64_t pj = mods[nQ + jj].q;
                y[jj] = nP == 1 ? cxP[0] : mul_shoup(cxP[jj * N], phat_inv[jj], phat_inv_s[jj], pj);
                if (rounded) v = __dadd_rn(v, __ddiv_rn((double)y[jj], (double)pj));
            }
        }
        const uint64_t r = rounded ? (uint64_t)__dadd_rn(v, 0.5) : 0;
        Split26 ys[A];
#pragma unroll
        for (int jj = 0; jj < A; jj++) ys[jj] = split26(P52 ? y[jj] : 0);
        for (size_t i = 0; i < Ql; i++) {
            const DevMod mq = mods[i];
            uint64_t val;
            if (P52 && nP > 1) {
                Acc52 acc{0, 0, 0};
                if (nP == (size_t)A) {
#pragma unroll
                    for (int jj = 0; jj < A; jj++) mac52(acc, ys[jj], phat_mod[i * nP + jj]);
                } else {
#pragma unroll
                    for (int jj = 0; jj < A; jj++)
                        if ((size_t)jj < nP) mac52(acc, ys[jj], phat_mod[i * nP + jj]);
                }
                if (rounded) mac52(acc, split26(mq.q - r), pmod[i]);
                val = reduce52(acc, mq);
            } else if (nP == 1) {
                val = reduce64(y[0], mq);
                if (r) val = sub_mod(val, reduce64(mods[nQ].q, mq), mq.q);
            } else {
                Acc128 acc{0, 0};
#pragma unroll
                for (int jj = 0; jj < A; jj++)
                    if ((size_t)jj < nP) mac128(acc, y[jj], phat_mod[i * nP + jj]);
                // - r P == (q - r) [P]_q (mod q), folded in before the one reduction
                if (rounded) mac128(acc, mq.q - r, pmod[i]);
                val = reduce128(acc, mq);
            }
            __builtin_nontemporal_store(val, &d[i * N]);
        }
    }
}

// delta[p][i][j], i < Ql-1: the exact centred conversion of polynomial p's
// coefficient-form limbs over B = {q_last, p_0 .. p_(nP-1)} (positions Ql-1 ..
// Ql+nP-1 of its QlP limbs) to q_i: sum_b y_b (B/b) - round(sum_b y_b / b) B with
// y_b = [x_b (B/b)^-1]_b, the representative of the value mod B in (-B/2, B/2],
// so that (x - delta) / B rounds x / (P q_last) to nearest. One thread per (p, j).
template <bool P52, int A>
__global__ void k_moddown_rescale_bconv(const uint64_t *__restrict__ cx, uint64_t *__restrict__ delta,
                                        const DevMod *__restrict__ mods, const uint64_t *__restrict__ bhat_inv,
                                        const uint64_t *__restrict__ bhat_inv_s,
                                        const uint64_t *__restrict__ bhat_mod, const uint64_t *__restrict__ bmod,
                                        size_t Ql, size_t nQ, size_t nP, size_t N, size_t npolys) {
    const size_t QlP = Ql + nP, L1 = Ql - 1, nb = nP + 1, total = npolys * N;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t p = t >> __builtin_ctzll(N), j = t & (N - 1);
        const uint64_t *xb = cx + (p * QlP + L1) * N + j;
        uint64_t *d = delta + p * L1 * N + j;
        uint64_t y[A + 1];
        double v = 0.0;
#pragma unroll
        for (int b = 0; b < A + 1; b++) {
            y[b] = 0;
            if ((size_t)b < nb) {
                const uint64_t qb = mods[b == 0 ? L1 : nQ + (b - 1)].q;
                y[b] = mul_shoup(xb[b * N], bhat_inv[b], bhat_inv_s[b], qb);
                v = __dadd_rn(v, __ddiv_rn((double)y[b], (double)qb));
            }
        }
        const uint64_t r = (uint64_t)__dadd_rn(v, 0.5);
        Split26 ys[A + 1];
#pragma unroll
        for (int b = 0; b < A + 1; b++) ys[b] = split26(P52 ? y[b] : 0);
        for (size_t i = 0; i < L1; i++) {
            const DevMod mq = mods[i];
            if (P52) {
                Acc52 acc{0, 0, 0};
                if (nb == (size_t)A + 1) {
#pragma unroll
                    for (int b = 0; b < A + 1; b++) mac52(acc, ys[b], bhat_mod[i * nb + b]);
                } else {
#pragma unroll
                    for (int b = 0; b < A + 1; b++)
                        if ((size_t)b < nb) mac52(acc, ys[b], bhat_mod[i * nb + b]);
                }
                mac52(acc, split26(mq.q - r), bmod[i]);
                __builtin_nontemporal_store(reduce52(acc, mq), &d[i * N]);
                continue;
            }
            Acc128 acc{0, 0};
#pragma unroll
            for (int b = 0; b < A + 1; b++)
                if ((size_t)b < nb) mac128(acc, y[b], bhat_mod[i * nb + b]);
            // - r B == (q - r) [B]_q (mod q), folded in before the one reduction
            mac128(acc, mq.q - r, bmod[i]);
            __builtin_nontemporal_store(reduce128(acc, mq), &d[i * N]);
        }
    }
}

// cx[m][0/1][k][j] = sum_g d[m][g][k][j] * key_g.(b/a)[limb(k)][j] for `members`
// digit sets (eval_key_switch.cu:12-60 per member).
// With perm, the digits are read through that NTT-domain Galois permutation
// (hoisted rotations of every member by one step).
// Layout: one thread per (member, limb k, coefficient j). A block covers
// J = 256 / MG consecutive coefficients and MG = 2^lgmg member slots (MG <= 16,
// so J >= 16 and the lanes of one member read whole 128-byte runs). Member slot 0
// loads the 2 beta key words of each coefficient once and stages them in LDS for
// every slot and member, so each key word leaves HBM once for the group and no
// thread holds the keys in registers (74 VGPRs at MAXB = 8 against 134, 5 waves
// per SIMD against 3). Every thread has only its own beta digit loads to wait for,
// and the grid holds members x QlP x N threads to hide their latency.
template <int MAXB, bool P52>
__global__ __launch_bounds__(256) void k_ks_inner_batch(const uint64_t *__restrict__ d, const uint64_t *__restrict__ key,
                                                        uint64_t *__restrict__ cx, const DevMod *__restrict__ mods,
                                                        size_t Ql, size_t nQ, size_t QP, size_t alpha, int beta,
                                                        size_t N, int members, int lgmg,
                                                        const uint32_t *__restrict__ perm, PolyPack c2, int own_c2,
                                                        int g0, int nb, int accumulate, PolyPack addp, int add_on,
                                                        const uint64_t *__restrict__ pmod,
                                                        const uint64_t *__restrict__ pmod_s) {
    const size_t QlP = Ql + alpha, total = QlP * N;
    const int lgj = 8 - lgmg;
    const int lane = (int)(threadIdx.x & ((1u << lgj) - 1));
    const size_t t = ((size_t)blockIdx.x << lgj) + (size_t)lane;
    const int ms = (int)(threadIdx.x >> lgj);
    const bool in = t < total;
    const size_t k = t >> __builtin_ctzll(N), j = t & (N - 1);
    const uint32_t pi = qlp_prime(k, Ql, nQ);
    // digits g0 .. g0 + nb - 1 of the beta of each member (nb <= MAXB). Member slot 0
    // loads the 2 nb key words of its coefficient once and leaves them in LDS for the
    // other slots: 16 bytes per digit and coefficient, both words split into 26-bit
    // halves (P52) or as they are.
    __shared__ uint4 kl[MAXB * 256];
    if (ms == 0 && in) {
        // every digit's two key words are requested before the first is used
        uint64_t wb[MAXB], wa[MAXB];
#pragma unroll
        for (int g = 0; g < MAXB; g++) {
            wb[g] = wa[g] = 0;
            if (g < nb) {
                const uint64_t *kg = key + (size_t)(g0 + g) * 2 * QP * N;
                wb[g] = kg[(size_t)pi * N + j];
                wa[g] = kg[(QP + pi) * N + j];
            }
        }
#pragma unroll
        for (int g = 0; g < MAXB; g++) {
            if (g < nb) {
                uint4 w;
                if (P52) {
                    const Split26 bs = split26(wb[g]), as = split26(wa[g]);
                    w = make_uint4(bs.lo, bs.hi, as.lo, as.hi);
                } else {
                    w = make_uint4((uint32_t)wb[g], (uint32_t)(wb[g] >> 32), (uint32_t)wa[g],
                                   (uint32_t)(wa[g] >> 32));
                }
                kl[g * 256 + lane] = w;
            }
        }
    }
    __syncthreads();
    if (!in || ms >= members) return;
    const size_t src = perm ? k * N + perm[j] : t;
    const DevMod m = mods[pi];
    // the digit among g0 .. g0 + nb - 1 whose own limbs hold k (digit dg's own limbs
    // k in [dg alpha, (dg + 1) alpha) are c2's limbs), or -1
    const int own_g = own_c2 && k < Ql ? (int)(k / alpha) - g0 : -1;
    // accumulate, or add P (d0, d1) on the data limbs (the merged ModDown + rescale
    // divides P d + cx by P q_last in one step), then store member mem's two outputs
    auto finish_member = [&](int mem, uint64_t v0, uint64_t v1) {
        uint64_t *o0 = cx + (size_t)(2 * mem) * total + t, *o1 = cx + (size_t)(2 * mem + 1) * total + t;
        if (accumulate) {
            v0 = add_mod(v0, *o0, m.q);
            v1 = add_mod(v1, *o1, m.q);
        } else if (add_on && k < Ql) {
            const uint64_t pm = pmod[pi], pms = pmod_s[pi];
            v0 = add_mod(v0, mul_shoup(addp.p[2 * mem][k * N + j], pm, pms, m.q), m.q);
            v1 = add_mod(v1, mul_shoup(addp.p[2 * mem + 1][k * N + j], pm, pms, m.q), m.q);
        }
        __builtin_nontemporal_store(v0, o0);
        __builtin_nontemporal_store(v1, o1);
    };
    // two members per step: their digit loads are in flight together
    const int mstep = 1 << lgmg;
    for (int mem = ms; mem < members; mem += 2 * mstep) {
        const int mem2 = mem + mstep;
        const bool two = mem2 < members;
        const int m2 = two ? mem2 : mem;
        const uint64_t *dm = d + ((size_t)mem * beta + g0) * total + src;
        const uint64_t *dm2 = d + ((size_t)m2 * beta + g0) * total + src;
        uint64_t dv[MAXB], dv2[MAXB];
#pragma unroll
        for (int g = 0; g < MAXB; g++) {
            dv[g] = g < nb ? (g == own_g ? c2.p[mem][src] : __builtin_nontemporal_load(&dm[(size_t)g * total])) : 0;
            dv2[g] = g < nb && two
                         ? (g == own_g ? c2.p[m2][src] : __builtin_nontemporal_load(&dm2[(size_t)g * total]))
                         : 0;
        }
        uint64_t v0, v1, w0, w1;
        if (P52) {
            Acc52 a0{0, 0, 0}, a1{0, 0, 0}, b0{0, 0, 0}, b1{0, 0, 0};
#pragma unroll
            for (int g = 0; g < MAXB; g++) {
                if (g < nb) {
                    // the key words are the pre-split operand, the digits are split once
                    const uint4 w = kl[g * 256 + lane];
                    const Split26 ds = split26(dv[g]), es = split26(dv2[g]);
                    a0.s0 += (uint64_t)ds.lo * w.x;
                    a0.s1 += (uint64_t)ds.hi * w.x;
                    a0.s1 += (uint64_t)ds.lo * w.y;
                    a0.s2 += (uint64_t)ds.hi * w.y;
                    a1.s0 += (uint64_t)ds.lo * w.z;
                    a1.s1 += (uint64_t)ds.hi * w.z;
                    a1.s1 += (uint64_t)ds.lo * w.w;
                    a1.s2 += (uint64_t)ds.hi * w.w;
                    b0.s0 += (uint64_t)es.lo * w.x;
                    b0.s1 += (uint64_t)es.hi * w.x;
                    b0.s1 += (uint64_t)es.lo * w.y;
                    b0.s2 += (uint64_t)es.hi * w.y;
                    b1.s0 += (uint64_t)es.lo * w.z;
                    b1.s1 += (uint64_t)es.hi * w.z;
                    b1.s1 += (uint64_t)es.lo * w.w;
                    b1.s2 += (uint64_t)es.hi * w.w;
                }
            }
            v0 = reduce52(a0, m);
            v1 = reduce52(a1, m);
            w0 = reduce52(b0, m);
            w1 = reduce52(b1, m);
        } else {
            Acc128 a0{0, 0}, a1{0, 0}, b0{0, 0}, b1{0, 0};
#pragma unroll
            for (int g = 0; g < MAXB; g++) {
                if (g < nb) {
                    const uint4 w = kl[g * 256 + lane];
                    const uint64_t kbw = ((uint64_t)w.y << 32) | w.x, kaw = ((uint64_t)w.w << 32) | w.z;
                    mac128(a0, dv[g], kbw);
                    mac128(a1, dv[g], kaw);
                    mac128(b0, dv2[g], kbw);
                    mac128(b1, dv2[g], kaw);
                }
            }
            v0 = reduce128(a0, m);
            v1 = reduce128(a1, m);
            w0 = reduce128(b0, m);
            w1 = reduce128(b1, m);
        }
        finish_member(mem, v0, v1);
        if (two) finish_member(mem2, w0, w1);
    }
}


// work counters per level (host side, relaxed): ModUp'd polynomials, inner-product
// output polynomials and ModDown'd polynomials
constexpr size_t kCounterLevels = 128;
std::atomic<uint64_t> g_cnt_modup[kCounterLevels], g_cnt_inner[kCounterLevels], g_cnt_moddown[kCounterLevels];

inline void bump(std::atomic<uint64_t> *a, size_t Ql, size_t by) {
    if (Ql < kCounterLevels) a[Ql].fetch_add(by, std::memory_order_relaxed);
}

}  // namespace

void ks_counters(uint64_t *modup, uint64_t *inner, uint64_t *moddown, size_t cap) {
    for (size_t i = 0; i < cap && i < kCounterLevels; i++) {
        if (modup) modup[i] = g_cnt_modup[i].load(std::memory_order_relaxed);
        if (inner) inner[i] = g_cnt_inner[i].load(std::memory_order_relaxed);
        if (moddown) moddown[i] = g_cnt_moddown[i].load(std::memory_order_relaxed);
    }
}

KsTuning &ks_tuning() {
    static KsTuning t;
    return t;
}

// Autotuner of the inner products' slot counts (the member / output grouping per
// block). The first call of a shape (kernel, N, Ql, members or outputs, digits,
// permuted, P d added, 52-bit primes) that writes its output outright times every
// candidate slot count on the call's own operands (a warm-up, then each candidate
// twice, the minimum kept) and caches the fastest; every candidate computes the
// same words, so the call's result is the same whichever ran last. Calls recorded
// into a graph, and accumulating calls, use the cached choice or one slot.
namespace {
struct TuneEntry {
    int best = 0;
    std::vector<float> us;
};
std::mutex g_tune_mu;
std::map<std::string, TuneEntry> g_tune;
std::atomic<bool> g_tune_on{true};

bool capturing(hipStream_t s) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &st) != hipSuccess) {
        (void)hipGetLastError();
        return true;
    }
    return st != hipStreamCaptureStatusNone;
}

// returns the slot count (log2) to use; `ran` tells whether the tuning runs
// already produced the call's output
int tuned_slots(const std::string &key, int max_lg, bool may_tune, hipStream_t s,
                const std::function<void(int)> &launch, bool *ran) {
    *ran = false;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune.find(key);
        if (it != g_tune.end()) return std::min(it->second.best, max_lg);
    }
    if (!may_tune || max_lg == 0 || !g_tune_on.load() || capturing(s)) return 0;
    const int nc = max_lg + 1;
    std::vector<hipEvent_t> ev(2 * nc * 2);
    for (auto &e : ev) SHM_HIP(hipEventCreate(&e));
    launch(0);  // warm-up
    for (int rep = 0; rep < 2; rep++)
        for (int lg = 0; lg < nc; lg++) {
            SHM_HIP(hipEventRecord(ev[2 * (rep * nc + lg)], s));
            launch(lg);
            SHM_HIP(hipEventRecord(ev[2 * (rep * nc + lg) + 1], s));
        }
    SHM_HIP(hipEventSynchronize(ev.back()));
    TuneEntry te;
    te.us.assign(nc, 1e30f);
    for (int rep = 0; rep < 2; rep++)
        for (int lg = 0; lg < nc; lg++) {
            float ms = 0.f;
            SHM_HIP(hipEventElapsedTime(&ms, ev[2 * (rep * nc + lg)], ev[2 * (rep * nc + lg) + 1]));
            te.us[lg] = std::min(te.us[lg], ms * 1000.f);
        }
    for (auto &e : ev) (void)hipEventDestroy(e);
    te.best = (int)(std::min_element(te.us.begin(), te.us.end()) - te.us.begin());
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune[key] = te;
    }
    *ran = true;
    return te.best;
}
}  // namespace

void ks_autotune_enable(bool on) { g_tune_on.store(on); }

void ks_autotune_clear() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune.clear();
}

std::string ks_autotune_report() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    std::string out;
    for (auto &[k, e] : g_tune) {
        out += k + " -> slots " + std::to_string(1 << e.best) + " (us:";
        for (size_t i = 0; i < e.us.size(); i++) {
            char b[32];
            snprintf(b, sizeof b, " %d:%.1f", 1 << i, e.us[i]);
            out += b;
        }
        out += ")\n";
    }
    return out;
}

void ks_counters_reset() {
    for (size_t i = 0; i < kCounterLevels; i++) {
        g_cnt_modup[i].store(0, std::memory_order_relaxed);
        g_cnt_inner[i].store(0, std::memory_order_relaxed);
        g_cnt_moddown[i].store(0, std::memory_order_relaxed);
    }
}

void ks_modup(const Context &c, const uint64_t *c2, size_t Ql, uint64_t *d, hipStream_t s, bool rounded) {
    const size_t N = c.N, alpha = c.nP, QlP = Ql + alpha;
    const LevelConsts &L = c.level(Ql);
    const size_t beta = L.beta;
    // the digits are multiplied by key digits 0 .. beta-1, and keys hold nQ / nP of
    // them: refuse before anything is launched (an inner product would read past
    // the key)
    if (beta > c.dnum) throw LogicError("key switching key has too few digits for this level");
    bump(g_cnt_modup, Ql, 1);
    Scratch tcoef(Ql * N * sizeof(uint64_t), s);
    PolyPack in{};
    in.p[0] = const_cast<uint64_t *>(c2);  // read only
    ntt_inverse_gather(c, in, tcoef.as<uint64_t>(), 1, Ql, 0, s);
    if (alpha > (size_t)kMaxAlpha) throw InvalidArgument("at most 16 special primes");
    with_alpha_bound(alpha, [&](auto bound) {
        constexpr int A = decltype(bound)::value;
        (c.primes52 ? k_modup_bconv<true, A> : k_modup_bconv<false, A>)<<<grid_for(beta * N), kBlock, 0, s>>>(
            tcoef.as<uint64_t>(), d, c.mods(), L.hat_inv.as<uint64_t>(), L.hat_inv_s.as<uint64_t>(),
            L.hat_mod.as<uint64_t>(), L.dig_mod.as<uint64_t>(), Ql, c.nQ, alpha, beta, N, 1, rounded ? 1 : 0, 1);
    });
    SHM_HIP(hipGetLastError());
    // one launch over every digit's QlP limbs (the own limbs were written in
    // coefficient form and transform back to their NTT-form input words)
    ntt_forward_qlp(c, d, Ql, 0, QlP, 0, 0, beta, QlP * N, s);
}

void ks_inner_multi(const Context &c, const uint64_t *d, size_t Ql, size_t ndig, const KeyRefs &kr, size_t nout,
                    bool local, uint64_t *cx, hipStream_t s, const uint32_t *perm, bool accumulate) {
    if (nout == 0 || nout > (size_t)kMaxKeys) throw InvalidArgument("ks_inner_multi: 1..34 outputs");
    if (ndig == 0) throw InvalidArgument("ks_inner_multi: no digits");
    const size_t QlP = Ql + c.nP, total = QlP * c.N;
    if (ndig > (size_t)kMaxDigits) {
        // digits in chunks of kMaxDigits: later chunks add their partial sums mod q
        for (size_t g0 = 0; g0 < ndig; g0 += kMaxDigits) {
            KeyRefs k2 = kr;
            for (size_t o = 0; o < nout; o++) k2.base[o] = kr.base[o] + g0 * kr.stride[o];
            ks_inner_multi(c, d + g0 * total, Ql, std::min<size_t>(kMaxDigits, ndig - g0), k2, nout, local, cx, s,
                           perm, accumulate || g0 > 0);
        }
        return;
    }
    bump(g_cnt_inner, Ql, nout);
    // output slots per block: the knob if set, else the autotuned choice (one slot,
    // each thread loading its digits once and looping over the outputs, until tuned)
    int max_lg = 0;
    while (max_lg < 4 && (size_t(1) << max_lg) < nout) max_lg++;
    auto fn = c.primes52
                  ? (ndig <= 4 ? k_ks_inner_multi<4, true> : ndig <= 8 ? k_ks_inner_multi<8, true> : k_ks_inner_multi<16, true>)
                  : (ndig <= 4 ? k_ks_inner_multi<4, false> : ndig <= 8 ? k_ks_inner_multi<8, false> : k_ks_inner_multi<16, false>);
    auto launch = [&](int lgog) {
        const unsigned blocks = (unsigned)((total + (256u >> lgog) - 1) >> (8 - lgog));
        fn<<<blocks, kBlock, 0, s>>>(d, kr, (int)nout, cx, c.mods(), Ql, c.nQ, c.nP, (int)ndig, c.N, local ? 1 : 0,
                                     perm, accumulate ? 1 : 0, lgog);
        SHM_HIP(hipGetLastError());
    };
    if (ks_tuning().multi_slots_log2 >= 0) {
        launch(std::min(ks_tuning().multi_slots_log2, max_lg));
        return;
    }
    const std::string tk = "multi N=" + std::to_string(c.N) + " Ql=" + std::to_string(Ql) +
                           " outputs=" + std::to_string(nout) + " digits=" + std::to_string(ndig) +
                           " perm=" + std::to_string(perm ? 1 : 0) + " local=" + std::to_string(local ? 1 : 0) +
                           " p52=" + std::to_string(c.primes52 ? 1 : 0);
    bool ran = false;
    const int lg = tuned_slots(tk, max_lg, !accumulate, s, launch, &ran);
    if (!ran) launch(lg);
}

void ks_moddown(const Context &c, uint64_t *cx, size_t npolys, size_t Ql, uint64_t *out, size_t out_stride,
                uint64_t acc_mask, hipStream_t s, bool rounded) {
    if (npolys == 0 || npolys > 64) throw InvalidArgument("ks_moddown: 1..64 polynomials");
    PolyPack pk{};
    for (size_t p = 0; p < npolys; p++) pk.p[p] = out + p * out_stride;
    ks_moddown_pack(c, cx, npolys, Ql, pk, acc_mask, s, rounded);
}

void ks_moddown_pack(const Context &c, uint64_t *cx, size_t npolys, size_t Ql, const PolyPack &out,
                     uint64_t acc_mask, hipStream_t s, bool rounded, const PolyPack *add_src, const uint32_t *perm) {
    const size_t N = c.N, nP = c.nP, QlP = Ql + nP;
    if (npolys == 0 || npolys > 64) throw InvalidArgument("ks_moddown: 1..64 polynomials");
    bump(g_cnt_moddown, Ql, npolys);
    Scratch delta(npolys * Ql * N * sizeof(uint64_t), s);
    ntt_inverse_qlp(c, cx, Ql, Ql, QlP, npolys, QlP * N, s);
    if (nP > (size_t)kMaxAlpha) throw InvalidArgument("at most 16 special primes");
    with_alpha_bound(nP, [&](auto bound) {
        constexpr int A = decltype(bound)::value;
        (c.primes52 ? k_moddown_bconv<true, A> : k_moddown_bconv<false, A>)<<<grid_for(npolys * N), kBlock, 0, s>>>(
            cx, delta.as<uint64_t>(), c.mods(), c.phat_inv.as<uint64_t>(), c.phat_inv_s.as<uint64_t>(),
            c.phat_mod.as<uint64_t>(), c.pmod.as<uint64_t>(), Ql, c.nQ, nP, N, npolys, rounded ? 1 : 0);
    });
    SHM_HIP(hipGetLastError());
    // the last NTT pass of delta writes (cx - delta) * P^-1 into the outputs
    moddown_ntt_fused(c, delta.as<uint64_t>(), cx, Ql, QlP, out, npolys, acc_mask, c.pinv.as<uint64_t>(),
                      c.pinv_s.as<uint64_t>(), s, add_src, perm);
}

void keyswitch(const Context &c, const uint64_t *c2, const uint64_t *key, size_t Ql, uint64_t *ct0, uint64_t *ct1,
               bool add_to_ct1, hipStream_t s, bool rounded) {
    const size_t N = c.N, alpha = c.nP, QlP = Ql + alpha;
    const size_t beta = c.level(Ql).beta;
    if (beta > c.dnum) throw LogicError("key switching key has too few digits for this level");
    Scratch d(beta * QlP * N * sizeof(uint64_t), s);
    Scratch cx(2 * QlP * N * sizeof(uint64_t), s);
    ks_modup(c, c2, Ql, d.as<uint64_t>(), s, rounded);
    bump(g_cnt_inner, Ql, 2);
    k_ks_inner<<<grid_for(QlP * N), kBlock, 0, s>>>(d.as<uint64_t>(), key, cx.as<uint64_t>(), c.mods(), Ql, c.nQ,
                                                    c.nQP(), alpha, beta, N);
    SHM_HIP(hipGetLastError());
    if (ct1 == ct0 + Ql * N) {
        ks_moddown(c, cx.as<uint64_t>(), 2, Ql, ct0, Ql * N, add_to_ct1 ? 3 : 1, s, rounded);
    } else {
        ks_moddown(c, cx.as<uint64_t>(), 1, Ql, ct0, Ql * N, 1, s, rounded);
        ks_moddown(c, cx.as<uint64_t>() + QlP * N, 1, Ql, ct1, Ql * N, add_to_ct1 ? 1 : 0, s, rounded);
    }
}

void ks_modup_batch(const Context &c, const PolyPack &c2, size_t m, size_t Ql, uint64_t *d, hipStream_t s,
                    bool rounded) {
    const size_t N = c.N, alpha = c.nP, QlP = Ql + alpha;
    const LevelConsts &L = c.level(Ql);
    const size_t beta = L.beta;
    if (m == 0 || m > (size_t)kPolyPack) throw InvalidArgument("ks_modup_batch: 1..64 members");
    if (beta > c.dnum) throw LogicError("key switching key has too few digits for this level");
    if (alpha > (size_t)kMaxAlpha) throw InvalidArgument("at most 16 special primes");
    bump(g_cnt_modup, Ql, m);
    // iNTT of the members into one buffer, conversion, NTT over m beta QlP limbs
    Scratch tcoef(m * Ql * N * sizeof(uint64_t), s);
    ntt_inverse_gather(c, c2, tcoef.as<uint64_t>(), m, Ql, 0, s);
    // From 4 members on, each digit gets its own NTT launch over the m members that
    // skips the digit's own limbs: they are c2's NTT-form words, which the inner
    // product reads from c2. With fewer members such launches are too small to
    // fill the GPU, so the own limbs are written in coefficient form and one launch
    // transforms every limb (back to c2's words for the own ones).
    const bool per_digit = m >= 4;
    with_alpha_bound(alpha, [&](auto bound) {
        constexpr int A = decltype(bound)::value;
        (c.primes52 ? k_modup_bconv<true, A> : k_modup_bconv<false, A>)<<<grid_for(m * beta * N), kBlock, 0, s>>>(
            tcoef.as<uint64_t>(), d, c.mods(), L.hat_inv.as<uint64_t>(), L.hat_inv_s.as<uint64_t>(),
            L.hat_mod.as<uint64_t>(), L.dig_mod.as<uint64_t>(), Ql, c.nQ, alpha, beta, N, m, rounded ? 1 : 0,
            per_digit ? 0 : 1);
    });
    SHM_HIP(hipGetLastError());
    if (per_digit) {
        for (size_t g = 0; g < beta; g++) {
            const size_t own_lo = g * alpha, own_hi = std::min(own_lo + alpha, Ql);
            ntt_forward_qlp(c, d + g * QlP * N, Ql, 0, QlP, own_lo, own_hi, m, beta * QlP * N, s);
        }
    } else {
        ntt_forward_qlp(c, d, Ql, 0, QlP, 0, 0, m * beta, QlP * N, s);
    }
}

void ks_inner_batch(const Context &c, const uint64_t *d, size_t m, size_t Ql, const uint64_t *key, uint64_t *cx,
                    hipStream_t s, const uint32_t *perm, const PolyPack *own, const PolyPack *add) {
    const size_t QlP = Ql + c.nP, beta = c.level(Ql).beta;
    if (beta > c.dnum) throw LogicError("key switching key has too few digits for this level");
    if (m == 0 || m > 64) throw InvalidArgument("ks_inner_batch: 1..64 members");
    bump(g_cnt_inner, Ql, 2 * m);
    // member slots per block: the knob if set, else the autotuned choice (one slot,
    // each thread loading its key words once and looping over the members, until tuned)
    int max_lg = 0;
    while (max_lg < 4 && (size_t(1) << max_lg) < m) max_lg++;
    const size_t total = QlP * c.N;
    // digits in chunks of kMaxDigits; a later chunk adds its partial sums mod q
    for (size_t g0 = 0; g0 < beta; g0 += kMaxDigits) {
        const size_t nb = std::min<size_t>(kMaxDigits, beta - g0);
        auto fn = c.primes52
                      ? (nb <= 4 ? k_ks_inner_batch<4, true> : nb <= 8 ? k_ks_inner_batch<8, true> : k_ks_inner_batch<16, true>)
                      : (nb <= 4 ? k_ks_inner_batch<4, false> : nb <= 8 ? k_ks_inner_batch<8, false> : k_ks_inner_batch<16, false>);
        auto launch = [&](int lgmg) {
            const unsigned blocks = (unsigned)((total + (256u >> lgmg) - 1) >> (8 - lgmg));
            fn<<<blocks, kBlock, 0, s>>>(d, key, cx, c.mods(), Ql, c.nQ, c.nQP(), c.nP, (int)beta, c.N, (int)m,
                                         lgmg, perm, own ? *own : PolyPack{}, own ? 1 : 0, (int)g0, (int)nb,
                                         g0 > 0 ? 1 : 0, add ? *add : PolyPack{}, add ? 1 : 0,
                                         c.pmod.as<uint64_t>(), c.pmod_s.as<uint64_t>());
            SHM_HIP(hipGetLastError());
        };
        if (ks_tuning().batch_slots_log2 >= 0) {
            launch(std::min(ks_tuning().batch_slots_log2, max_lg));
            continue;
        }
        const std::string tk = "batch N=" + std::to_string(c.N) + " Ql=" + std::to_string(Ql) +
                               " m=" + std::to_string(m) + " digits=" + std::to_string(nb) +
                               " perm=" + std::to_string(perm ? 1 : 0) + " add=" + std::to_string(add ? 1 : 0) +
                               " p52=" + std::to_string(c.primes52 ? 1 : 0);
        bool ran = false;
        const int lg = tuned_slots(tk, max_lg, g0 == 0, s, launch, &ran);
        if (!ran) launch(lg);
    }
}

void keyswitch_batch(const Context &c, const PolyPack &c2, size_t m, const uint64_t *key, size_t Ql,
                     const PolyPack &ct01, uint64_t acc_mask, hipStream_t s, bool rounded) {
    const size_t N = c.N, QlP = Ql + c.nP, beta = c.level(Ql).beta;
    if (m == 0 || 2 * m > (size_t)kPolyPack) throw InvalidArgument("keyswitch_batch: 1..32 members");
    Scratch d(m * beta * QlP * N * sizeof(uint64_t), s);
    ks_modup_batch(c, c2, m, Ql, d.as<uint64_t>(), s, rounded);
    // inner products: each key word read once for the group
    Scratch cx(2 * m * QlP * N * sizeof(uint64_t), s);
    ks_inner_batch(c, d.as<uint64_t>(), m, Ql, key, cx.as<uint64_t>(), s, nullptr, &c2);
    // ModDown of the 2m polynomials into the members' (c0, c1)
    ks_moddown_pack(c, cx.as<uint64_t>(), 2 * m, Ql, ct01, acc_mask, s, rounded);
}

void keyswitch_rescale_batch(const Context &c, const PolyPack &c2, const PolyPack &d01, size_t m, const uint64_t *key,
                             size_t Ql, const PolyPack &out, hipStream_t s) {
    const size_t N = c.N, nP = c.nP, QlP = Ql + nP, beta = c.level(Ql).beta;
    if (m == 0 || 2 * m > (size_t)kPolyPack) throw InvalidArgument("keyswitch_rescale_batch: 1..32 members");
    if (Ql < 2) throw LogicError("end of modulus switching chain reached");
    if (nP > (size_t)kMaxAlpha) throw InvalidArgument("at most 16 special primes");
    const LevelConsts &L = c.level(Ql);
    Scratch d(m * beta * QlP * N * sizeof(uint64_t), s);
    ks_modup_batch(c, c2, m, Ql, d.as<uint64_t>(), s, true);
    // cx = P (d0, d1) + the inner products: the value the merged step divides
    Scratch cx(2 * m * QlP * N * sizeof(uint64_t), s);
    ks_inner_batch(c, d.as<uint64_t>(), m, Ql, key, cx.as<uint64_t>(), s, nullptr, &c2, &d01);
    // B = {q_last, P}: inverse NTT of those limbs, exact centred conversion to the
    // other Ql - 1 primes, and the NTT of the correction whose last pass writes
    // (cx - delta) B^-1 into the outputs
    bump(g_cnt_moddown, Ql, 2 * m);
    ntt_inverse_qlp(c, cx.as<uint64_t>(), Ql, Ql - 1, QlP, 2 * m, QlP * N, s);
    Scratch delta(2 * m * (Ql - 1) * N * sizeof(uint64_t), s);
    with_alpha_bound(nP, [&](auto bound) {
        constexpr int A = decltype(bound)::value;
        (c.primes52 ? k_moddown_rescale_bconv<true, A>
                    : k_moddown_rescale_bconv<false, A>)<<<grid_for(2 * m * N), kBlock, 0, s>>>(
            cx.as<uint64_t>(), delta.as<uint64_t>(), c.mods(), L.bhat_inv.as<uint64_t>(),
            L.bhat_inv_s.as<uint64_t>(), L.bhat_mod.as<uint64_t>(), L.bmod.as<uint64_t>(), Ql, c.nQ, nP, N, 2 * m);
    });
    SHM_HIP(hipGetLastError());
    moddown_ntt_fused(c, delta.as<uint64_t>(), cx.as<uint64_t>(), Ql - 1, QlP, out, 2 * m, 0,
                      L.binv.as<uint64_t>(), L.binv_s.as<uint64_t>(), s);
}

}  // namespace shm
