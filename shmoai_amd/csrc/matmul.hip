// MOAI column-packed ciphertext x plaintext matmul
// (Ct_pt_matrix_mul.cuh:18-185 ct_pt_matrix_mul_wo_pre):
//   out_i = rescale( sum_{j < row_W} X_j (.) encode(W[j][i], X_j.scale) ),
//   out_i.scale = X_0.scale.
// encode() of a scalar fills every slot, so the plaintext is the constant
// polynomial round(fl(W[j][i] * scale_j)) and, in NTT form, every slot holds
// its residue: the contraction is, per limb l and coefficient column c,
//   acc_l[i][c] = sum_j [w_ji]_{q_l} * X_j[l][c]  (mod q_l).
//
// Variants:
//   MATMUL_PER_CT  - the reference's structure: for each (i, j) one constant
//                    plaintext, one multiply_plain and one add over the whole
//                    ciphertext, then a rescale per output.
//   MATMUL_U64     - one batched contraction kernel (64-bit modular MACs with
//                    128-bit accumulation) and one batched rescale.
//   MATMUL_INT8    - the same contraction on INT8 MFMA: X limbs split into
//                    8-bit digits, weights into balanced signed 8-bit digits,
//                    exact int32 accumulation, recombined with 128-bit integer
//                    arithmetic and one Barrett reduction (bit-identical).
// Both batched variants contract at most 256 rows of W per launch; longer
// contractions (row_W = 768 or 3072, the BERT projection and FFN shapes of
// ct_pt_matrix_mul_wo_pre_large, Ct_pt_matrix_mul.cuh:584-700, which computes
// the same sums) run in row chunks whose reduced partial sums are added mod q.
#include <cmath>
#include <vector>

#include "engine.h"

namespace shm {

namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(size_t work) {
    size_t g = (work + kBlock - 1) / kBlock;
    if (g > (1u << 20)) g = 1u << 20;
    return (unsigned)(g ? g : 1);
}

float g_last_ms = 0.f;
float g_last_tile_ms = 0.f;

// Weight residues: wres[l][j][i] = [round(fl(W[j][i] * scale_j))]_{q_l}.
__global__ void k_weight_residues(const double *W, const double *scales, uint64_t *wres, const DevMod *mods,
                                  size_t row_w, size_t col_w, size_t Ql) {
    const size_t total = Ql * row_w * col_w;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t l = t / (row_w * col_w), r = t % (row_w * col_w), j = r / col_w, i = r % col_w;
        const double v = round(W[j * col_w + i] * scales[j]);
        const DevMod m = mods[l];
        const bool neg = signbit(v);
        const double a = fabs(v);
        uint64_t res;
        if (a < 18446744073709551616.0) {
            res = reduce64((uint64_t)a, m);
        } else {
            const uint64_t lo = (uint64_t)fmod(a, 18446744073709551616.0);
            const uint64_t hi = (uint64_t)(a / 18446744073709551616.0);
            res = barrett128(hi, lo, m);
        }
        if (neg && res) res = m.q - res;
        wres[t] = res;
    }
}

// Fused contraction (64-bit VALU). Block = (limb l, poly p, 32-column tile);
// X tile staged in LDS as [j][32]; 8 groups of 32 lanes each own col_w/8
// outputs, 128-bit accumulators, one Barrett reduction per output.
template <int OUT_PER_THREAD>
__global__ __launch_bounds__(256) void k_contract_u64(const uint64_t *const *__restrict__ X,
                                                      const uint64_t *__restrict__ wres,
                                                      uint64_t *__restrict__ acc_out, const DevMod *__restrict__ mods,
                                                      size_t row_w, size_t col_w, size_t Ql, size_t N,
                                                      int accumulate) {
    constexpr int TN = 32;
    extern __shared__ uint64_t xs[];  // [row_w][TN]
    const size_t tiles = N / TN;
    const size_t b = blockIdx.x;
    const size_t tile = b % tiles, lp = b / tiles;  // lp = p * Ql + l
    const size_t l = lp % Ql;
    const size_t off = lp * N + tile * TN;
    for (size_t t = threadIdx.x; t < row_w * TN; t += blockDim.x) {
        const size_t j = t / TN, n = t % TN;
        xs[t] = X[j][off + n];
    }
    __syncthreads();
    const DevMod m = mods[l];
    const int n = threadIdx.x % TN;
    const int grp = threadIdx.x / TN;  // 8 groups
    const uint64_t *wl = wres + l * row_w * col_w;
    for (size_t i0 = (size_t)grp * OUT_PER_THREAD; i0 < col_w; i0 += 8 * OUT_PER_THREAD) {
        Acc128 acc[OUT_PER_THREAD];
#pragma unroll
        for (int k = 0; k < OUT_PER_THREAD; k++) acc[k] = Acc128{0, 0};
        for (size_t j = 0; j < row_w; j++) {
            const uint64_t x = xs[j * TN + n];
#pragma unroll
            for (int k = 0; k < OUT_PER_THREAD; k++)
                if (i0 + k < col_w) mac128(acc[k], x, wl[j * col_w + i0 + k]);
        }
#pragma unroll
        for (int k = 0; k < OUT_PER_THREAD; k++) {
            if (i0 + k >= col_w) continue;
            uint64_t *dst = &acc_out[(i0 + k) * 2 * Ql * N + off + n];
            uint64_t v = reduce128(acc[k], m);
            // a later row chunk adds to the sum of the earlier ones
            *dst = accumulate ? add_mod(v, *dst, m.q) : v;
        }
    }
}

// acc += x * w (constant) for the per-ciphertext reference structure.
__global__ void k_mul_const_add(uint64_t *acc, const uint64_t *x, const uint64_t *pt, const DevMod *mods, size_t Ql,
                                size_t N, size_t total, int first) {
    const size_t S = Ql * N;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t r = t % S;
        const DevMod m = mods[r / N];
        const uint64_t v = mul_mod(x[t], pt[r], m);
        acc[t] = first ? v : add_mod(acc[t], v, m.q);
    }
}

// Masked form: up to kMaskJ input ciphertexts of one row block, by value.
constexpr int kMaskJ = 8;
struct XBlock {
    const uint64_t *p[kMaskJ];
};

// acc[cc][poly][t] (+)= sum_{jj < jb} X_jj[poly][t] * pt[jj * cb + cc][t] for
// t < Ql N: the products of one block of inputs and one block of outputs, with
// 128-bit sums and one Barrett reduction per output word; `first` starts the
// sums, otherwise they are added mod q to the earlier row blocks'.
__global__ void k_mask_mac(XBlock xb, int jb, const uint64_t *__restrict__ pt, int cb, uint64_t *__restrict__ acc,
                           const DevMod *__restrict__ mods, size_t Ql, size_t N, int first) {
    const size_t S = Ql * N;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < S; t += (size_t)gridDim.x * blockDim.x) {
        const DevMod m = mods[t >> __builtin_ctzll(N)];
        uint64_t x0[kMaskJ], x1[kMaskJ];
#pragma unroll
        for (int jj = 0; jj < kMaskJ; jj++) {
            x0[jj] = jj < jb ? xb.p[jj][t] : 0;
            x1[jj] = jj < jb ? xb.p[jj][S + t] : 0;
        }
        for (int cc = 0; cc < cb; cc++) {
            Acc128 a0{0, 0}, a1{0, 0};
#pragma unroll
            for (int jj = 0; jj < kMaskJ; jj++) {
                if (jj < jb) {
                    const uint64_t w = pt[(size_t)(jj * cb + cc) * S + t];
                    mac128(a0, x0[jj], w);
                    mac128(a1, x1[jj], w);
                }
            }
            uint64_t v0 = reduce128(a0, m), v1 = reduce128(a1, m);
            uint64_t *o = acc + (size_t)cc * 2 * S + t;
            if (!first) {
                v0 = add_mod(v0, o[0], m.q);
                v1 = add_mod(v1, o[S], m.q);
            }
            o[0] = v0;
            o[S] = v1;
        }
    }
}

// w[b] = W[j0 + jj][c0 + cc], fix[b] = fixj[j0 + jj] for b = jj * cb + cc.
__global__ void k_mask_weights(const double *W, const double *fixj, double *w, double *fix, size_t j0, size_t c0,
                               int jb, int cb, size_t col_w) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= jb * cb) return;
    const int jj = b / cb, cc = b % cb;
    w[b] = W[(j0 + jj) * col_w + c0 + cc];
    fix[b] = fixj[j0 + jj];
}

// Batched rescale helpers over many ciphertexts (each [2][Ql][N]).
__global__ void k_gather_last(const uint64_t *acc, uint64_t *last, size_t Ql, size_t N, size_t npolys) {
    const size_t total = npolys * N;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t p = t >> __builtin_ctzll(N), j = t & (N - 1);
        last[t] = acc[(p * Ql + Ql - 1) * N + j];
    }
}

void batched_rescale(const Context &c, const uint64_t *acc, const std::vector<Ciphertext *> &out, size_t Ql,
                     hipStream_t s) {
    const size_t N = c.N, L1 = Ql - 1, nct = out.size(), np = 2 * nct;
    Scratch last(np * N * sizeof(uint64_t), s);
    Scratch delta(np * L1 * N * sizeof(uint64_t), s);
    k_gather_last<<<grid_for(np * N), kBlock, 0, s>>>(acc, last.as<uint64_t>(), Ql, N, np);
    ntt_inverse(c, last.as<uint64_t>(), np, 1, L1, s);
    std::vector<uint64_t *> ptrs(nct);
    for (size_t i = 0; i < nct; i++) ptrs[i] = out[i]->data();
    Scratch dptrs(nct * sizeof(uint64_t *), s);
    SHM_HIP(hipMemcpyAsync(dptrs.as<void>(), ptrs.data(), nct * sizeof(uint64_t *), hipMemcpyHostToDevice, s));
    const LevelConsts &L = c.level(Ql);
    rescale_ntt_fused(c, last.as<uint64_t>(), acc, Ql, delta.as<uint64_t>(), dptrs.as<uint64_t *>(), nullptr, 2, np,
                      L1, L.qlast_inv.as<uint64_t>(), L.qlast_inv_s.as<uint64_t>(), s);
    SHM_HIP(hipGetLastError());
    // the host vector of pointers must outlive the async copy
    SHM_HIP(hipStreamSynchronize(s));
}

}  // namespace

void ctpt_matmul_int8(const Context &c, const std::vector<const Ciphertext *> &X, const double *W, size_t row_w,
                      size_t col_w, uint64_t *acc, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1,
                      const Int8WindowHook *hook = nullptr);
bool int8_supported(const Context &c, size_t Ql, const double *W, const double *scales, size_t row_w, size_t col_w,
                    int *dx_out, int *dw_out);

int g_last_variant = -1;

void ctpt_matmul_masked(const Context &c, const std::vector<const Ciphertext *> &X, const double *W, size_t row_w,
                        size_t col_w, const uint8_t *mask, const std::vector<Ciphertext *> &out, hipStream_t s) {
    if (X.size() < row_w || out.size() != col_w || row_w == 0 || col_w == 0)
        throw InvalidArgument("bad dimensions of X or W");
    const size_t Ql = X[0]->Ql(), N = c.N, S = Ql * N, n = N / 2;
    if (Ql < 2) throw LogicError("end of modulus switching chain reached");
    for (size_t j = 0; j < row_w; j++) {
        if (X[j]->chain != X[0]->chain || X[j]->size != 2)
            throw InvalidArgument("encrypted and plain parameter mismatch");
        if (!(X[j]->scale > 0)) throw InvalidArgument("scale out of bounds");
    }
    const double out_scale = X[0]->scale;
    for (Ciphertext *o : out) o->resize(2, X[0]->chain + 1, s);
    // plaintexts of one (row block, output block) are encoded, transformed and
    // contracted together
    const size_t JB = std::min<size_t>(row_w, kMaskJ), CB = std::min<size_t>(col_w, 16);
    std::vector<double> fixj(row_w);
    for (size_t j = 0; j < row_w; j++) fixj[j] = X[j]->scale / (double)n;  // encode()'s scale / n
    Scratch acc(col_w * 2 * S * sizeof(uint64_t), s);
    Scratch pt(JB * CB * S * sizeof(uint64_t), s);
    Scratch fft(JB * CB * n * sizeof(double2), s);
    Scratch dmask(n, s);
    Scratch dW(row_w * col_w * sizeof(double), s);
    Scratch dfixj(row_w * sizeof(double), s);
    Scratch dw(JB * CB * sizeof(double), s);
    Scratch dfix(JB * CB * sizeof(double), s);
    SHM_HIP(hipMemcpyAsync(dmask.as<void>(), mask, n, hipMemcpyHostToDevice, s));
    SHM_HIP(hipMemcpyAsync(dW.as<void>(), W, row_w * col_w * sizeof(double), hipMemcpyHostToDevice, s));
    SHM_HIP(hipMemcpyAsync(dfixj.as<void>(), fixj.data(), row_w * sizeof(double), hipMemcpyHostToDevice, s));
    for (size_t c0 = 0; c0 < col_w; c0 += CB) {
        const int cb = (int)std::min(CB, col_w - c0);
        for (size_t j0 = 0; j0 < row_w; j0 += JB) {
            const int jb = (int)std::min(JB, row_w - j0);
            k_mask_weights<<<1, kBlock, 0, s>>>(dW.as<double>(), dfixj.as<double>(), dw.as<double>(),
                                                 dfix.as<double>(), j0, c0, jb, cb, col_w);
            encode_masked_batch(c, dw.as<double>(), dfix.as<double>(), dmask.as<uint8_t>(), (size_t)(jb * cb), Ql,
                                pt.as<uint64_t>(), fft.as<double2>(), s);
            XBlock xb{};
            for (int jj = 0; jj < jb; jj++) xb.p[jj] = X[j0 + jj]->data();
            k_mask_mac<<<grid_for(S), kBlock, 0, s>>>(xb, jb, pt.as<uint64_t>(), cb, acc.as<uint64_t>() + c0 * 2 * S,
                                                      c.mods(), Ql, N, j0 == 0 ? 1 : 0);
            SHM_HIP(hipGetLastError());
        }
    }
    batched_rescale(c, acc.as<uint64_t>(), out, Ql, s);
    SHM_HIP(hipStreamSynchronize(s));
    for (Ciphertext *o : out) o->scale = out_scale;
}

float last_matmul_kernel_ms() { return g_last_ms; }
float last_matmul_rescale_tile_ms() { return g_last_tile_ms; }
int last_matmul_variant() { return g_last_variant; }

void ctpt_matmul(const Context &c, const std::vector<const Ciphertext *> &X, const double *W, size_t row_w,
                 size_t col_w, const std::vector<Ciphertext *> &out, int variant, hipStream_t s) {
    if (X.size() < row_w || out.size() != col_w || row_w == 0 || col_w == 0)
        throw InvalidArgument("bad dimensions of X or W");
    const size_t Ql = X[0]->Ql(), N = c.N, S = Ql * N;
    if (Ql < 2) throw LogicError("end of modulus switching chain reached");
    for (size_t j = 0; j < row_w; j++) {
        if (X[j]->chain != X[0]->chain || X[j]->size != 2)
            throw InvalidArgument("encrypted and plain parameter mismatch");
    }
    const double out_scale = X[0]->scale;
    for (Ciphertext *o : out) o->resize(2, X[0]->chain + 1, s);

    std::vector<double> scales(row_w);
    for (size_t j = 0; j < row_w; j++) scales[j] = X[j]->scale;
    if (variant == 2) {
        // shapes, primes or weights outside the INT8 digit bounds run the 64-bit contraction
        int dx = 0, dw = 0;
        if (!int8_supported(c, Ql, W, scales.data(), row_w, col_w, &dx, &dw)) variant = 0;
    }
    g_last_variant = variant;
    hipEvent_t e0, e1;
    SHM_HIP(hipEventCreate(&e0));
    SHM_HIP(hipEventCreate(&e1));

    if (variant == 1) {
        // Reference structure: per output, per input: encode scalar, multiply_plain, add; rescale.
        Scratch acc(2 * S * sizeof(uint64_t), s);
        Scratch pt(S * sizeof(uint64_t), s);
        SHM_HIP(hipEventRecord(e0, s));
        for (size_t i = 0; i < col_w; i++) {
            for (size_t j = 0; j < row_w; j++) {
                encode_scalar(c, W[j * col_w + i], Ql, scales[j], pt.as<uint64_t>(), s);
                k_mul_const_add<<<grid_for(2 * S), kBlock, 0, s>>>(acc.as<uint64_t>(), X[j]->data(),
                                                                   pt.as<uint64_t>(), c.mods(), Ql, N, 2 * S,
                                                                   j == 0 ? 1 : 0);
            }
            launch_rescale(c, acc.as<uint64_t>(), out[i]->data(), Ql, 2, s);
        }
        SHM_HIP(hipEventRecord(e1, s));
    } else {
        Scratch acc(col_w * 2 * S * sizeof(uint64_t), s);
        if (variant == 2 && row_w <= 256 && ks_tuning().matmul_overlap) {
            // Overlap: the last limb is contracted first and the context's second stream
            // starts the rescale's correction from it (gather, inverse NTT); the other
            // limbs are contracted in windows on s while the second stream rescales each
            // finished window (the INT8 contraction is MFMA work, the rescale NTT VALU
            // work). X is read once; the outputs equal the serial path's.
            Context &cc = const_cast<Context &>(c);
            const hipStream_t r = cc.aux_stream();
            const size_t L1 = Ql - 1, nct = out.size(), np = 2 * nct;
            std::vector<uint64_t *> ptrs(nct);
            for (size_t i = 0; i < nct; i++) ptrs[i] = out[i]->data();
            Scratch dptrs(nct * sizeof(uint64_t *), s);
            SHM_HIP(hipMemcpyAsync(dptrs.as<void>(), ptrs.data(), nct * sizeof(uint64_t *), hipMemcpyHostToDevice, s));
            Scratch last(np * N * sizeof(uint64_t), r);
            Scratch delta(np * L1 * N * sizeof(uint64_t), r);
            const LevelConsts &L = c.level(Ql);
            std::vector<hipEvent_t> evs;
            auto handoff = [&] {
                hipEvent_t ev;
                SHM_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
                evs.push_back(ev);
                SHM_HIP(hipEventRecord(ev, s));
                SHM_HIP(hipStreamWaitEvent(r, ev, 0));
            };
            std::vector<std::pair<hipEvent_t, hipEvent_t>> tile_ev;
            Int8WindowHook hook;
            hook.window = (L1 + 2) / 3;
            hook.last_done = [&] {
                handoff();
                k_gather_last<<<grid_for(np * N), kBlock, 0, r>>>(acc.as<uint64_t>(), last.as<uint64_t>(), Ql, N, np);
                SHM_HIP(hipGetLastError());
                ntt_inverse(c, last.as<uint64_t>(), np, 1, L1, r);
            };
            hook.window_done = [&](size_t lo, size_t n) {
                handoff();
                hipEvent_t t0, t1;
                SHM_HIP(hipEventCreate(&t0));
                SHM_HIP(hipEventCreate(&t1));
                tile_ev.emplace_back(t0, t1);
                set_rescale_tile_events(t0, t1);
                rescale_ntt_fused_window(c, last.as<uint64_t>(), acc.as<uint64_t>(), Ql, delta.as<uint64_t>(),
                                         dptrs.as<uint64_t *>(), 2, np, L1, lo, n, L.qlast_inv.as<uint64_t>(),
                                         L.qlast_inv_s.as<uint64_t>(), r);
                set_rescale_tile_events(nullptr, nullptr);
            };
            try {
                ctpt_matmul_int8(c, X, W, row_w, col_w, acc.as<uint64_t>(), s, e0, e1, &hook);
            } catch (...) {
                set_rescale_tile_events(nullptr, nullptr);
                (void)hipStreamSynchronize(r);
                throw;
            }
            // s (which frees acc, and owns the outputs) continues after the rescale
            hipEvent_t done;
            SHM_HIP(hipEventCreateWithFlags(&done, hipEventDisableTiming));
            SHM_HIP(hipEventRecord(done, r));
            SHM_HIP(hipStreamWaitEvent(s, done, 0));
            SHM_HIP(hipEventSynchronize(done));
            SHM_HIP(hipEventRecord(e1, s));  // the step: contraction and rescale together
            g_last_tile_ms = 0.f;
            for (auto &[t0, t1] : tile_ev) {
                float ms = 0.f;
                SHM_HIP(hipEventElapsedTime(&ms, t0, t1));
                g_last_tile_ms += ms;
                (void)hipEventDestroy(t0);
                (void)hipEventDestroy(t1);
            }
            (void)hipEventDestroy(done);
            for (hipEvent_t ev : evs) (void)hipEventDestroy(ev);
        } else if (variant == 2) {
            ctpt_matmul_int8(c, X, W, row_w, col_w, acc.as<uint64_t>(), s, e0, e1);
        } else {
            // X rows staged in LDS per launch: at most 256 rows (64 KiB); longer
            // contractions run in balanced row chunks added mod q
            constexpr size_t kRows = 256;
            const size_t nch = (row_w + kRows - 1) / kRows, kc_max = (row_w + nch - 1) / nch;
            Scratch dW(row_w * col_w * sizeof(double), s);
            Scratch dscale(row_w * sizeof(double), s);
            Scratch wres(Ql * kc_max * col_w * sizeof(uint64_t), s);
            Scratch xptr(row_w * sizeof(uint64_t *), s);
            std::vector<const uint64_t *> xp(row_w);
            for (size_t j = 0; j < row_w; j++) xp[j] = X[j]->data();
            SHM_HIP(hipMemcpyAsync(dW.as<void>(), W, row_w * col_w * sizeof(double), hipMemcpyHostToDevice, s));
            SHM_HIP(hipMemcpyAsync(dscale.as<void>(), scales.data(), row_w * sizeof(double), hipMemcpyHostToDevice, s));
            SHM_HIP(hipMemcpyAsync(xptr.as<void>(), xp.data(), row_w * sizeof(uint64_t *), hipMemcpyHostToDevice, s));
            const size_t blocks = 2 * Ql * (N / 32);
            size_t k0 = 0;
            for (size_t ch = 0; ch < nch; ch++) {
                const size_t kc = row_w / nch + (ch < row_w % nch ? 1 : 0);
                k_weight_residues<<<grid_for(Ql * kc * col_w), kBlock, 0, s>>>(
                    dW.as<double>() + k0 * col_w, dscale.as<double>() + k0, wres.as<uint64_t>(), c.mods(), kc, col_w,
                    Ql);
                const size_t lds = kc * 32 * sizeof(uint64_t);
                if (ch == 0) SHM_HIP(hipEventRecord(e0, s));
                k_contract_u64<8><<<(unsigned)blocks, 256, lds, s>>>(xptr.as<const uint64_t *>() + k0,
                                                                     wres.as<uint64_t>(), acc.as<uint64_t>(),
                                                                     c.mods(), kc, col_w, Ql, N, ch > 0 ? 1 : 0);
                SHM_HIP(hipGetLastError());
                k0 += kc;
            }
            SHM_HIP(hipEventRecord(e1, s));
            SHM_HIP(hipStreamSynchronize(s));
        }
        if (!(variant == 2 && row_w <= 256 && ks_tuning().matmul_overlap)) {
            hipEvent_t t0, t1;
            SHM_HIP(hipEventCreate(&t0));
            SHM_HIP(hipEventCreate(&t1));
            set_rescale_tile_events(t0, t1);
            try {
                batched_rescale(c, acc.as<uint64_t>(), out, Ql, s);
            } catch (...) {
                set_rescale_tile_events(nullptr, nullptr);
                throw;
            }
            set_rescale_tile_events(nullptr, nullptr);
            SHM_HIP(hipEventSynchronize(t1));
            SHM_HIP(hipEventElapsedTime(&g_last_tile_ms, t0, t1));
            SHM_HIP(hipEventDestroy(t0));
            SHM_HIP(hipEventDestroy(t1));
        }
    }
    SHM_HIP(hipEventSynchronize(e1));
    SHM_HIP(hipEventElapsedTime(&g_last_ms, e0, e1));
    SHM_HIP(hipEventDestroy(e0));
    SHM_HIP(hipEventDestroy(e1));
    for (Ciphertext *o : out) o->scale = out_scale;
}

}  // namespace shm
