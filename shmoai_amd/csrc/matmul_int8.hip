// INT8-MFMA contraction for the MOAI ct-pt matmul (bit-identical to the
// 64-bit path).
//
// Per limb l (prime q) and coefficient column c:
//   Y[i][c] = sum_j w_ij X_j[c]  (mod q),  w_ij = round(fl(W[j][i] * scale_j))
// (X_j (.) [w]_q == X_j * w mod q, so the signed integer weight is shared by
// all limbs). Split
//   X = sum_a 256^a (x'_a + 128),  x'_a in [-128, 127]   (DX unsigned digits)
//   w = sum_b 256^b w_b,           w_b in [-128, 127]    (DW balanced digits)
// so Y = sum_d 256^d C_d + 128 * R * rowsum_i with C_d = sum_{a+b=d} W_b X'_a
// and R = sum_a 256^a. Each W_b X'_a is a 16x16x64 int8 MFMA GEMM
// accumulating straight into the int32 plane C_{a+b} (|C_d| < 2^24, exact).
// The planes are recombined by Horner's rule in signed 128-bit integers and
// reduced once with Barrett; the rowsum correction is added mod q.
//
// Mapping: block = 4 waves = 64 coefficient columns of one (poly, limb); a
// wave owns 16 columns and keeps their X digits as MFMA B fragments in
// registers (DX * KS fragments), then sweeps the output rows 16 at a time,
// streaming the W digit fragments (A operand) from L2.
#include <cmath>
#include <vector>

#include "engine.h"

namespace shm {

namespace {

typedef int8_t i8x16 __attribute__((ext_vector_type(16)));
typedef int32_t i32x4 __attribute__((ext_vector_type(4)));

constexpr int kBlock = 256;

inline unsigned grid_for(size_t work) {
    size_t g = (work + kBlock - 1) / kBlock;
    if (g > (1u << 20)) g = 1u << 20;
    return (unsigned)(g ? g : 1);
}

__device__ __forceinline__ int64_t weight_int(const double *W, const double *scales, size_t col_w, size_t j,
                                              size_t i) {
    return (int64_t)round(W[j * col_w + i] * scales[j]);
}

// A fragments in phase-major order, so that the fragments of one row phase
// (RB blocks of 16 output rows) are one contiguous run:
// frag[(((ph * DW + b) * RB + ibl) * KS + ks) * 64 + lane] = 16 bytes,
// byte e = digit b of w(row m = 16 (ph RB + ibl) + lane % 16,
//                       k = 64 ks + 16 (lane / 16) + e); zero outside the matrix.
__global__ void k_weight_frags(const double *W, const double *scales, i8x16 *frag, size_t row_w, size_t col_w,
                               int DW, int MB, int KS, int RB) {
    const int NPH = (MB + RB - 1) / RB;
    const size_t total = (size_t)NPH * DW * RB * KS * 64;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int lane = (int)(t % 64);
        size_t r = t / 64;
        const int ks = (int)(r % KS);
        r /= KS;
        const int ibl = (int)(r % RB);
        r /= RB;
        const int b = (int)(r % DW);
        const int ph = (int)(r / DW);
        const size_t m = 16 * ((size_t)ph * RB + ibl) + lane % 16;
        i8x16 v;
#pragma unroll
        for (int e = 0; e < 16; e++) {
            const size_t k = 64 * (size_t)ks + 16 * (lane / 16) + e;
            int8_t d = 0;
            if (m < col_w && k < row_w) {
                int64_t w = weight_int(W, scales, col_w, k, m);
                int64_t dig = 0;
                for (int bb = 0; bb <= b; bb++) {
                    dig = w & 255;
                    if (dig >= 128) dig -= 256;
                    w = (w - dig) >> 8;
                }
                d = (int8_t)dig;
            }
            v[e] = d;
        }
        frag[t] = v;
    }
}

// corr[l][i] = [128 R rowsum_i - 2^106]_q (the 2^106 offset keeps the
// recombined plane sum non-negative); mu[l] = floor(2^109 / q_l).
__global__ void k_weight_corr(const double *W, const double *scales, uint64_t *corr, uint64_t *mu,
                              const DevMod *mods, size_t row_w, size_t col_w, size_t Ql, int DX) {
    const size_t total = Ql * col_w;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t l = t / col_w, i = t % col_w;
        const DevMod m = mods[l];
        int64_t s = 0;
        for (size_t j = 0; j < row_w; j++) s += weight_int(W, scales, col_w, j, i);
        uint64_t rs = reduce64((uint64_t)(s < 0 ? -s : s), m);
        if (s < 0 && rs) rs = m.q - rs;
        uint64_t R = 0;
        for (int a = 0; a < DX; a++) R = add_mod(mul_mod(R, 256, m), 1, m.q);
        const uint64_t c = mul_mod(128 % m.q, R, m);
        uint64_t p106 = 1;
        for (int e = 0; e < 106; e++) p106 = add_mod(p106, p106, m.q);
        corr[t] = sub_mod(mul_mod(c, rs, m), p106, m.q);
        if (i == 0) {
            // floor(2^109 / q) from floor(2^128 / q): shift by 19 bits (fits 64 bits for q > 2^45)
            mu[l] = (m.mu_hi << 45) | (m.mu_lo >> 19);
        }
    }
}

__device__ __forceinline__ uint32_t bperm(uint32_t hi_src, uint32_t lo_src, uint32_t sel) {
    return __builtin_amdgcn_perm(hi_src, lo_src, sel);
}

// 4x4 byte transpose: o_k = (w0.byte k, w1.byte k, w2.byte k, w3.byte k).
__device__ __forceinline__ void transpose4(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3, uint32_t (&o)[4]) {
    const uint32_t t0 = bperm(w1, w0, 0x06020400u);  // w0.0 w1.0 w0.2 w1.2
    const uint32_t t1 = bperm(w1, w0, 0x07030501u);  // w0.1 w1.1 w0.3 w1.3
    const uint32_t t2 = bperm(w3, w2, 0x06020400u);  // w2.0 w3.0 w2.2 w3.2
    const uint32_t t3 = bperm(w3, w2, 0x07030501u);  // w2.1 w3.1 w2.3 w3.3
    o[0] = bperm(t2, t0, 0x05040100u);
    o[1] = bperm(t3, t1, 0x05040100u);
    o[2] = bperm(t2, t0, 0x07060302u);
    o[3] = bperm(t3, t1, 0x07060302u);
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Registers: the X digit fragments (DX KS 4 VGPRs) and the 11 planes dominate.
// Up to DX KS = 21 (e.g. 7 digits x 192 rows) the kernel is held to 3 waves per
// SIMD (156 registers at 7 x 3, no spill), which measured faster than
// prefetching the next phase's weight fragments at 2 waves per SIMD; larger
// shapes get 2 waves per SIMD so that they do not spill.
template <int DX, int DW, int KS, int RB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(DX * KS <= 21 ? 3 : 2))) void k_contract_int8(const uint64_t *const *__restrict__ X,
                                                       const i8x16 *__restrict__ frag,
                                                       const uint64_t *__restrict__ corr,
                                                       const uint64_t *__restrict__ mu_arr,
                                                       uint64_t *__restrict__ acc_out,
                                                       const DevMod *__restrict__ mods, int row_w, int col_w,
                                                       int MB, size_t Ql, size_t N, int accumulate, int radix,
                                                       int l_lo, int l_n) {
    constexpr int ND = DX + DW - 1;
    // The planes are recombined modulo 2^128, so S = sum_k x'_k w_k comes out exactly as
    // long as its true value satisfies |S| < 2^106 (x' the centred digits, |x'| < 2^(8 DX - 1)),
    // which the host checks from the weights before launching.
    static_assert(ND >= 9 && ND <= 12, "plane recombination covers 9 to 12 planes");
    static_assert(DX <= 8, "at most 8 digits per limb");
    constexpr int FR = DW * RB * KS * 64;  // fragments staged per row phase
    __shared__ i8x16 wl[FR];
    __shared__ uint64_t cl[RB * 16];  // row-sum corrections of this limb for the phase's rows
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const size_t tiles = N / 64;
    // the launch covers limbs l_lo .. l_lo + l_n - 1 of both polynomials
    const size_t tile = blockIdx.x % tiles, lw = blockIdx.x / tiles;
    const size_t l = (size_t)l_lo + lw % (size_t)l_n, lp = (lw / (size_t)l_n) * Ql + l;  // lp = p * Ql + l
    const size_t col = tile * 64 + 16 * wave + (lane & 15);
    const size_t off = lp * N + col;
    const uint64_t q = mods[l].q, mu = mu_arr[l];
    // row pointers staged in LDS once per block; rows k >= row_w point at row 0
    // so that every load is unconditional (their words are zeroed below)
    __shared__ const uint64_t *xl[KS * 64];
    for (int t = threadIdx.x; t < KS * 64; t += 256) xl[t] = X[t < row_w ? t : 0];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");

    // B fragments: lane holds X'_a[k = 64 ks + 16 (lane/16) + e][col] for
    // e = 0..15, built by byte transposes of the little-endian limb words.
    u32x4 B[DX][KS];
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
        uint64_t x[16];
#pragma unroll
        for (int e = 0; e < 16; e++) {
            const int k = 64 * ks + 16 * (lane >> 4) + e;
            const uint64_t v = xl[k][off];
            x[e] = k < row_w ? v : 0;
        }
#pragma unroll
        for (int g = 0; g < 4; g++) {
            uint32_t lo[4], hi[4];
            transpose4((uint32_t)x[4 * g], (uint32_t)x[4 * g + 1], (uint32_t)x[4 * g + 2], (uint32_t)x[4 * g + 3], lo);
            transpose4((uint32_t)(x[4 * g] >> 32), (uint32_t)(x[4 * g + 1] >> 32), (uint32_t)(x[4 * g + 2] >> 32),
                       (uint32_t)(x[4 * g + 3] >> 32), hi);
#pragma unroll
            for (int a = 0; a < DX; a++) B[a][ks][g] = (a < 4 ? lo[a & 3] : hi[a & 3]) ^ 0x80808080u;
        }
    }
    // rows k >= row_w carry digit -128 in B; their A entries are 0 so they add nothing.

    // Block barrier that orders LDS only: a full __syncthreads would also wait
    // for this wave's outstanding output stores before every phase.
    auto lds_barrier = [] {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    };
    // 2^{8e} from the runtime radix (256): kept out of constant folding so that each
    // term is one signed 32 x 32 + 64 multiply-add instead of shifts and moves
    int32_t pw[4];
    pw[0] = 1;
#pragma unroll
    for (int e = 1; e < 4; e++) pw[e] = pw[e - 1] * radix;
    const int nph = (MB + RB - 1) / RB;
    for (int ph = 0; ph < nph; ph++) {
        lds_barrier();
        const i8x16 *src = frag + (size_t)ph * FR;
        for (int t = threadIdx.x; t < FR; t += 256) wl[t] = src[t];
        for (int t = threadIdx.x; t < RB * 16; t += 256) {
            const int row = 16 * RB * ph + t;
            cl[t] = row < col_w ? corr[l * col_w + row] : 0;
        }
        lds_barrier();
#pragma unroll 1
        for (int ibl = 0; ibl < RB; ibl++) {
            const int ib = ph * RB + ibl;
            if (ib >= MB) break;
            i32x4 C[ND];
#pragma unroll
            for (int d = 0; d < ND; d++) C[d] = i32x4{0, 0, 0, 0};
#pragma unroll
            for (int b = 0; b < DW; b++) {
#pragma unroll
                for (int ks = 0; ks < KS; ks++) {
                    const i8x16 A = wl[((b * RB + ibl) * KS + ks) * 64 + lane];
#pragma unroll
                    for (int a = 0; a < DX; a++)
                        C[a + b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A, __builtin_bit_cast(i8x16, B[a][ks]),
                                                                         C[a + b], 0, 0, 0);
                }
            }
            // lane holds rows 16 ib + 4 (lane/16) + r, column col
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = 16 * ib + 4 * (lane >> 4) + r;
                // S = sum_d C_d 256^d as three signed 64-bit sums of 32 x 32 products,
                // g_k = sum_{e<4} C_{4k+e} 2^{8e}, S = g0 + g1 2^32 + g2 2^64
                int64_t g0 = 0, g1 = 0, g2 = 0;
#pragma unroll
                for (int d = 0; d < ND; d++) {
                    const int64_t t = (int64_t)C[d][r] * (int64_t)pw[d & 3];
                    if (d < 4) g0 += t; else if (d < 8) g1 += t; else g2 += t;
                }
                // T = S + 2^106 (0 < T < 2^107)
                const uint64_t a_lo = (uint64_t)g1 << 32;
                const uint64_t lo = (uint64_t)g0 + a_lo;
                const uint64_t hi = (uint64_t)(g0 >> 63) + (uint64_t)(g1 >> 32) + (lo < a_lo ? 1 : 0) + (uint64_t)g2 +
                                    (uint64_t(1) << 42);
                // quotient from the top 62 bits: T' = T >> 45, qe = floor(T' mu / 2^64)
                const uint64_t tp = (hi << 19) | (lo >> 45);
                const uint64_t qe = __umul64hi(tp, mu);
                uint64_t res = lo - qe * q;  // T - qe q < 4q
                res = res >= 2 * q ? res - 2 * q : res;
                res = res >= q ? res - q : res;
                if (row < col_w) {
                    res = add_mod(res, cl[row - 16 * RB * ph], q);
                    uint64_t *dst = &acc_out[(size_t)row * 2 * Ql * N + off];
                    // a later row chunk adds to the sum of the earlier ones
                    if (accumulate) res = add_mod(res, __builtin_nontemporal_load(dst), q);
                    // streamed once by the rescale that follows: keep it out of the caches
                    __builtin_nontemporal_store(res, dst);
                }
            }
        }
    }
}

typedef void (*ContractFn)(const uint64_t *const *, const i8x16 *, const uint64_t *, const uint64_t *, uint64_t *,
                           const DevMod *, int, int, int, size_t, size_t, int, int, int, int);

constexpr int rb_for(int KS) { return KS == 1 ? 4 : KS == 2 ? 4 : KS == 3 ? 3 : 2; }

ContractFn pick(int DX, int DW, int KS) {
#define SHM_PICK(dx, dw)                                                           \
    if (DX == dx && DW == dw) {                                                    \
        if (KS == 1) return k_contract_int8<dx, dw, 1, rb_for(1)>;                 \
        if (KS == 2) return k_contract_int8<dx, dw, 2, rb_for(2)>;                 \
        if (KS == 3) return k_contract_int8<dx, dw, 3, rb_for(3)>;                 \
        if (KS == 4) return k_contract_int8<dx, dw, 4, rb_for(4)>;                 \
    }
    SHM_PICK(6, 4)
    SHM_PICK(6, 5)
    SHM_PICK(6, 6)
    SHM_PICK(7, 4)
    SHM_PICK(7, 5)
    SHM_PICK(7, 6)
    SHM_PICK(8, 4)
#undef SHM_PICK
    return nullptr;
}

}  // namespace

// Rows of W contracted by one launch. A launch keeps the X digit fragments of
// all its rows in registers (at most 4 x 64 rows); longer contractions run in
// row chunks whose reduced sums are added mod q into the accumulators.
constexpr size_t kChunkRows = 256;

// Returns false when the primes or the weights fall outside the INT8 path.
bool int8_supported(const Context &c, size_t Ql, const double *W, const double *scales, size_t row_w, size_t col_w,
                    int *dx_out, int *dw_out) {
    if (c.N % 64) return false;
    int qbits = 0, qmin = 64;
    for (size_t l = 0; l < Ql; l++) {
        int b = 64 - __builtin_clzll(c.q[l]);
        qbits = std::max(qbits, b);
        qmin = std::min(qmin, b);
    }
    // The single-word quotient estimate qe = floor(floor(T / 2^45) mu / 2^64), with
    // mu = floor(2^109 / q) and 0 < T < 2^107, satisfies 0 <= T - qe q < q (1 + T / 2^109 + 2^45 / q)
    // < 2.25 q for 2^45 < q (bit length >= 46), which two conditional subtractions reduce,
    // and mu fits 64 bits for q > 2^45; q < 2^62 keeps the byte split within 8 digits.
    if (qmin < 46 || qbits > 61) return false;
    const int DX = (qbits + 7) / 8;
    double mx = 0;
    for (size_t j = 0; j < row_w; j++)
        for (size_t i = 0; i < col_w; i++) mx = std::max(mx, std::fabs(std::round(W[j * col_w + i] * scales[j])));
    int DW = 1;
    // balanced digits: D digits cover |w| <= 127 * (256^D - 1) / 255
    while (DW < 9 && mx > 127.0 * (std::pow(256.0, DW) - 1) / 255.0) DW++;
    if (DW < 4) DW = 4;
    // the recombined sum S = sum_k x'_k w_k of a row chunk (x' the centred digits of a
    // limb, |x'| < 2^(8 DX - 1)) must stay below 2^106 in magnitude
    const size_t nch = (row_w + kChunkRows - 1) / kChunkRows, kc_max = (row_w + nch - 1) / nch;
    if ((double)kc_max * std::ldexp(1.0, 8 * DX - 1) * mx >= std::ldexp(1.0, 106)) return false;
    if (!pick(DX, DW, 1)) return false;
    *dx_out = DX;
    *dw_out = DW;
    return true;
}

void ctpt_matmul_int8(const Context &c, const std::vector<const Ciphertext *> &X, const double *W, size_t row_w,
                      size_t col_w, uint64_t *acc, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1,
                      const Int8WindowHook *hook) {
    const size_t Ql = X[0]->Ql(), N = c.N;
    std::vector<double> scales(row_w);
    for (size_t j = 0; j < row_w; j++) scales[j] = X[j]->scale;
    int DX = 0, DW = 0;
    if (!int8_supported(c, Ql, W, scales.data(), row_w, col_w, &DX, &DW))
        throw InvalidArgument("weights or primes outside the INT8 MFMA contraction");
    // balanced row chunks of at most kChunkRows
    const size_t nch = (row_w + kChunkRows - 1) / kChunkRows;
    const size_t kc_max = (row_w + nch - 1) / nch;
    if (hook && nch != 1) throw LogicError("limb windows need one row chunk");
    const int MB = (int)((col_w + 15) / 16);
    auto frag_count = [&](int KS) {
        const int RB = rb_for(KS), NPH = (MB + RB - 1) / RB;
        return (size_t)NPH * DW * RB * KS * 64;
    };
    Scratch dW(row_w * col_w * sizeof(double), s);
    Scratch dscale(row_w * sizeof(double), s);
    Scratch frag(frag_count((int)((kc_max + 63) / 64)) * sizeof(i8x16), s);
    Scratch corr(Ql * col_w * sizeof(uint64_t), s);
    Scratch mu(Ql * sizeof(uint64_t), s);
    Scratch xptr(row_w * sizeof(uint64_t *), s);
    std::vector<const uint64_t *> xp(row_w);
    for (size_t j = 0; j < row_w; j++) xp[j] = X[j]->data();
    SHM_HIP(hipMemcpyAsync(dW.as<void>(), W, row_w * col_w * sizeof(double), hipMemcpyHostToDevice, s));
    SHM_HIP(hipMemcpyAsync(dscale.as<void>(), scales.data(), row_w * sizeof(double), hipMemcpyHostToDevice, s));
    SHM_HIP(hipMemcpyAsync(xptr.as<void>(), xp.data(), row_w * sizeof(uint64_t *), hipMemcpyHostToDevice, s));
    const size_t tiles = N / 64;
    size_t k0 = 0;
    for (size_t ch = 0; ch < nch; ch++) {
        // chunk rows [k0, k0 + kc): W rows are contiguous, so the chunk is a sub-matrix
        const size_t kc = row_w / nch + (ch < row_w % nch ? 1 : 0);
        const int KS = (int)((kc + 63) / 64), RB = rb_for(KS);
        const size_t nfrag = frag_count(KS);
        const double *Wc = dW.as<double>() + k0 * col_w;
        const double *sc = dscale.as<double>() + k0;
        k_weight_frags<<<grid_for(nfrag), kBlock, 0, s>>>(Wc, sc, frag.as<i8x16>(), kc, col_w, DW, MB, KS, RB);
        k_weight_corr<<<grid_for(Ql * col_w), kBlock, 0, s>>>(Wc, sc, corr.as<uint64_t>(), mu.as<uint64_t>(),
                                                              c.mods(), kc, col_w, Ql, DX);
        SHM_HIP(hipGetLastError());
        ContractFn fn = pick(DX, DW, KS);
        auto launch = [&](size_t l_lo, size_t l_n) {
            hipLaunchKernelGGL(fn, dim3((unsigned)(2 * l_n * tiles)), dim3(256), 0, s,
                               xptr.as<const uint64_t *>() + k0, frag.as<i8x16>(), corr.as<uint64_t>(),
                               mu.as<uint64_t>(), acc, c.mods(), (int)kc, (int)col_w, MB, Ql, N, ch > 0 ? 1 : 0, 256,
                               (int)l_lo, (int)l_n);
            SHM_HIP(hipGetLastError());
        };
        if (ch == 0) SHM_HIP(hipEventRecord(ev0, s));
        if (hook) {
            // the last limb first (the rescale's correction needs only it), then the
            // other limbs in windows, each handed to the hook once it is on the stream
            launch(Ql - 1, 1);
            hook->last_done();
            for (size_t w0 = 0; w0 < Ql - 1; w0 += hook->window) {
                const size_t wn = std::min(hook->window, Ql - 1 - w0);
                launch(w0, wn);
                hook->window_done(w0, wn);
            }
        } else {
            launch(0, Ql);
        }
        k0 += kc;
    }
    SHM_HIP(hipEventRecord(ev1, s));
    SHM_HIP(hipStreamSynchronize(s));
}

}  // namespace shm
