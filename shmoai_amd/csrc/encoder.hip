// CKKS canonical-embedding encoder/decoder (FP64), on device.
//
// Encode (ckks.cu:89-170): bit-reverse the slot vector, special inverse FFT
// (fft.cu:226-356, Gentleman-Sande stages with twiddle roots[M - psiIdx]) with
// the last stage scaled by scale/n, round half away from zero, reduce per
// prime (negative v -> q - (|v| mod q)), place coefficient ci at ci*N/(2n),
// then NTT. Decode (ckks.cu:172-246) is iNTT, CRT composition to a centred
// multi-precision integer, conversion to double by 64-bit words, the forward
// special FFT (fft.cu:88-224) and a bit reversal.
//
// This file is compiled with -ffp-contract=off: every product and sum is
// rounded separately, in the operation order written here.
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

__device__ __forceinline__ uint32_t brv_bits(uint32_t x, int bits) { return bits ? (__brev(x) >> (32 - bits)) : 0; }

__global__ void k_bitrev_load(const double *re, const double *im, size_t nv, double2 *buf, uint32_t n, int logn) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        double2 v = make_double2(0.0, 0.0);
        if (t < nv) v = make_double2(re[t], im ? im[t] : 0.0);
        buf[brv_bits(t, logn)] = v;
    }
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    const double t0 = a.x * b.x;
    const double t1 = a.y * b.y;
    const double t2 = a.x * b.y;
    const double t3 = a.y * b.x;
    return make_double2(t0 - t1, t2 + t3);
}

__device__ __forceinline__ uint32_t psi_index(const uint32_t *group, uint32_t k, int logPairs, int logn, int logRatio,
                                              uint32_t M) {
    const uint32_t bk = logn > 1 ? brv_bits(k << logPairs, logn - 1) : 0;
    uint32_t idx = group[bk];
    idx = (uint32_t)(((uint64_t)idx << (logPairs + logRatio)) & (M - 1));
    return idx;
}

// Butterfly tid of one Gentleman-Sande stage of the inverse special FFT.
__device__ __forceinline__ void ifft_butterfly(double2 *buf, const double2 *roots, const uint32_t *group,
                                               uint32_t tid, int logn, int logPairs, int logRatio, uint32_t M,
                                               double scalar, int last) {
    const uint32_t pairs = 1u << logPairs;
    const uint32_t k = tid >> logPairs, j = tid & (pairs - 1);
    const uint32_t g = 2 * k * pairs + j;
    const double2 tw = roots[M - psi_index(group, k, logPairs, logn, logRatio, M)];
    const double2 a = buf[g], b = buf[g + pairs];
    double2 s0 = make_double2(a.x + b.x, a.y + b.y);
    double2 s1 = make_double2(a.x - b.x, a.y - b.y);
    s1 = cmul(s1, tw);
    if (last && scalar != 0.0) {
        s0 = make_double2(s0.x * scalar, s0.y * scalar);
        s1 = make_double2(s1.x * scalar, s1.y * scalar);
    }
    buf[g] = s0;
    buf[g + pairs] = s1;
}

// One Gentleman-Sande stage of the inverse special FFT.
__global__ void k_ifft_stage(double2 *buf, const double2 *roots, const uint32_t *group, uint32_t n, int logn,
                             int logPairs, int logRatio, uint32_t M, double scalar, int last) {
    for (uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x; tid < n / 2; tid += gridDim.x * blockDim.x)
        ifft_butterfly(buf, roots, group, tid, logn, logPairs, logRatio, M, scalar, last);
}

// The same stage over a batch of vectors: buf[b][n], last-stage scalar fix[b].
__global__ void k_ifft_stage_batch(double2 *buf, const double2 *roots, const uint32_t *group, uint32_t n, int logn,
                                   int logPairs, int logRatio, uint32_t M, const double *fix, int last) {
    const uint32_t b = blockIdx.y;
    for (uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x; tid < n / 2; tid += gridDim.x * blockDim.x)
        ifft_butterfly(buf + (size_t)b * n, roots, group, tid, logn, logPairs, logRatio, M, fix[b], last);
}

// Slot vectors w[b] * mask, bit-reversed: exactly the values a host vector
// {mask[t] == 1 ? w : 0.0} loads (Ct_pt_matrix_mul.cuh:931-938).
__global__ void k_mask_load_batch(const double *w, const uint8_t *mask, double2 *buf, uint32_t n, int logn) {
    const uint32_t b = blockIdx.y;
    const double v = w[b];
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x)
        buf[(size_t)b * n + brv_bits(t, logn)] = make_double2(mask[t] ? v : 0.0, 0.0);
}

// One Cooley-Tukey stage of the forward special FFT.
__global__ void k_fft_stage(double2 *buf, const double2 *roots, const uint32_t *group, uint32_t n, int logn,
                            int logPairs, int logRatio, uint32_t M) {
    const uint32_t pairs = 1u << logPairs;
    for (uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x; tid < n / 2; tid += gridDim.x * blockDim.x) {
        const uint32_t k = tid >> logPairs, j = tid & (pairs - 1);
        const uint32_t g = 2 * k * pairs + j;
        const double2 tw = roots[psi_index(group, k, logPairs, logn, logRatio, M)];
        const double2 s0 = buf[g];
        const double2 s1 = cmul(buf[g + pairs], tw);
        buf[g] = make_double2(s0.x + s1.x, s0.y + s1.y);
        buf[g + pairs] = make_double2(s0.x - s1.x, s0.y - s1.y);
    }
}

// rns_base.cu:51-130 decompose_array (pt zero-filled beforehand).
__global__ void k_decompose(const double2 *buf, uint64_t *pt, const DevMod *mods, uint32_t n, size_t ratio, size_t Ql,
                            size_t N) {
    const size_t total = Ql * 2 * (size_t)n;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t k = t / (2 * n), ci = t % (2 * n);
        const double v = ci < n ? round(buf[ci].x) : round(buf[ci - n].y);
        pt[k * N + ci * ratio] = residue_of_rounded(v, mods[k]);
    }
}

// Constant plaintext in NTT form: every slot holds [round(fl(value*scale))]_q.
__global__ void k_fill_scalar(uint64_t *pt, const DevMod *mods, double v, size_t N, size_t total) {
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const DevMod m = mods[t >> __builtin_ctzll(N)];
        uint64_t r = residue_of_rounded(v, m);
        pt[t] = r >= m.q ? r - m.q : r;
    }
}

// rns_base.cu:200-262 compose_kernel: CRT compose, centre, convert.
// qhat: [Ql][W] words, Q: [W], half: [W]. W = Ql <= 64 words.
__global__ void k_compose(const uint64_t *x, double2 *buf, const DevMod *mods, const uint64_t *qhat,
                          const uint64_t *qhinv, const uint64_t *qhinv_s, const uint64_t *Q, const uint64_t *half,
                          size_t Ql, size_t N, uint32_t n, size_t ratio, double inv_scale) {
    for (size_t ci = blockIdx.x * (size_t)blockDim.x + threadIdx.x; ci < 2 * (size_t)n;
         ci += (size_t)gridDim.x * blockDim.x) {
        uint64_t acc[64];
        const size_t W = Ql;
        for (size_t w = 0; w < W; w++) acc[w] = 0;
        for (size_t i = 0; i < Ql; i++) {
            const uint64_t q = mods[i].q;
            const uint64_t prod = mul_shoup(x[i * N + ci * ratio], qhinv[i], qhinv_s[i], q);
            // acc += qhat_i * prod (multi-precision), then reduce mod Q
            uint64_t carry = 0;
            for (size_t w = 0; w < W; w++) {
                const uint64_t a = qhat[i * W + w];
                const uint64_t lo = a * prod, hi = __umul64hi(a, prod);
                uint64_t s = acc[w] + lo;
                uint64_t c1 = s < lo;
                s += carry;
                c1 += s < carry;
                acc[w] = s;
                carry = hi + c1;
            }
            bool ge = carry != 0;
            if (!ge) {
                ge = true;
                for (size_t w = W; w-- > 0;) {
                    if (acc[w] != Q[w]) {
                        ge = acc[w] > Q[w];
                        break;
                    }
                }
            }
            if (ge) {
                uint64_t borrow = 0;
                for (size_t w = 0; w < W; w++) {
                    const uint64_t a = acc[w], b = Q[w];
                    const uint64_t d = a - b - borrow;
                    borrow = (a < b) || (a == b && borrow) ? 1 : 0;
                    acc[w] = d;
                }
            }
        }
        bool upper = true;
        for (size_t w = W; w-- > 0;) {
            if (acc[w] != half[w]) {
                upper = acc[w] > half[w];
                break;
            }
        }
        double res = 0.0, sc = inv_scale;
        const double two64 = 18446744073709551616.0;
        if (upper) {
            for (size_t w = 0; w < W; w++, sc *= two64) {
                if (acc[w] > Q[w]) {
                    const uint64_t d = acc[w] - Q[w];
                    res += d ? (double)d * sc : 0.0;
                } else {
                    const uint64_t d = Q[w] - acc[w];
                    res -= d ? (double)d * sc : 0.0;
                }
            }
        } else {
            for (size_t w = 0; w < W; w++, sc *= two64) {
                const uint64_t d = acc[w];
                res += d ? (double)d * sc : 0.0;
            }
        }
        if (ci < n)
            buf[ci].x = res;
        else
            buf[ci - n].y = res;
    }
}

__global__ void k_bitrev_store(const double2 *buf, double *re, double *im, uint32_t n, int logn) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const double2 v = buf[t];
        const uint32_t r = brv_bits(t, logn);
        re[r] = v.x;
        im[r] = v.y;
    }
}

int ilog2(size_t v) {
    int l = 0;
    while ((size_t(1) << l) < v) l++;
    return l;
}

typedef unsigned __int128 hu128;

// Every slot holds one complex constant (the vector(slot_count, value) the
// facade's encode(double) builds, ckks_evaluator_parallel.cuh:73-78); a constant
// vector is its own bit reversal.
__global__ void k_fill_const(double2 *buf, double2 v, uint32_t n) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) buf[t] = v;
}

// batch_input's interleaved column packing (Batch_encode_encrypt.cuh:26-33): plaintext
// b holds column i = col0 + b with slot num_x * k + j = X[j][k][i] (X is
// [num_x][num_row][num_col]) and zeros beyond num_x * num_row, bit-reversed.
__global__ void k_columns_load_batch(const double *X, uint32_t num_x, uint32_t num_row, uint32_t num_col,
                                     uint32_t col0, double2 *buf, uint32_t n, int logn) {
    const uint32_t b = blockIdx.y, i = col0 + b, used = num_x * num_row;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        double v = 0.0;
        if (t < used) {
            const uint32_t j = t % num_x, k = t / num_x;
            v = X[((size_t)j * num_row + k) * num_col + i];
        }
        buf[(size_t)b * n + brv_bits(t, logn)] = make_double2(v, 0.0);
    }
}

// The inverse special FFT of a loaded, bit-reversed slot vector (last stage scaled
// by scale / n), rounding and reduction per prime, and the forward NTT.
void encode_finish(const Context &c, double2 *buf, size_t n, int logn, int logRatio, size_t M, size_t Ql,
                   double scale, uint64_t *out, hipStream_t s) {
    const size_t N = c.N;
    const double fix = scale / (double)n;
    for (int iter = logn - 1; iter >= 0; iter--) {
        const int logPairs = logn - iter - 1;
        const int last = (size_t(1) << logPairs) == n / 2;
        k_ifft_stage<<<grid_for(n / 2), kBlock, 0, s>>>(buf, c.enc_roots.as<double2>(), c.enc_group.as<uint32_t>(),
                                                        (uint32_t)n, logn, logPairs, logRatio, (uint32_t)M, fix, last);
    }
    launch_fill_zero(out, Ql * N, s);
    k_decompose<<<grid_for(Ql * 2 * n), kBlock, 0, s>>>(buf, out, c.mods(), (uint32_t)n, N / (2 * n), Ql, N);
    SHM_HIP(hipGetLastError());
    ntt_forward(c, out, 1, Ql, 0, s);
}

}  // namespace

void encode(const Context &c, const double *re, const double *im, size_t n_values, size_t Ql, double scale,
            uint64_t *out, hipStream_t s) {
    const size_t N = c.N, n = N / 2, M = 2 * N;
    if (n_values > n) throw InvalidArgument("values_size is too large");
    if (!(scale > 0)) throw InvalidArgument("scale out of bounds");
    const int logn = ilog2(n), logRatio = ilog2(M) - logn - 2;
    Scratch vals(2 * std::max<size_t>(n_values, 1) * sizeof(double), s);
    Scratch buf(n * sizeof(double2), s);
    double *dre = vals.as<double>(), *dim = dre + std::max<size_t>(n_values, 1);
    if (n_values) {
        SHM_HIP(hipMemcpyAsync(dre, re, n_values * sizeof(double), hipMemcpyHostToDevice, s));
        if (im) SHM_HIP(hipMemcpyAsync(dim, im, n_values * sizeof(double), hipMemcpyHostToDevice, s));
    }
    k_bitrev_load<<<grid_for(n), kBlock, 0, s>>>(dre, im ? dim : nullptr, n_values, buf.as<double2>(), (uint32_t)n,
                                                 logn);
    encode_finish(c, buf.as<double2>(), n, logn, logRatio, M, Ql, scale, out, s);
}

void encode_const(const Context &c, double re, double im, size_t Ql, double scale, uint64_t *out, hipStream_t s) {
    const size_t N = c.N, n = N / 2, M = 2 * N;
    if (!(scale > 0)) throw InvalidArgument("scale out of bounds");
    const int logn = ilog2(n), logRatio = ilog2(M) - logn - 2;
    Scratch buf(n * sizeof(double2), s);
    k_fill_const<<<grid_for(n), kBlock, 0, s>>>(buf.as<double2>(), make_double2(re, im), (uint32_t)n);
    encode_finish(c, buf.as<double2>(), n, logn, logRatio, M, Ql, scale, out, s);
}

void encode_columns_batch(const Context &c, const double *X, size_t num_x, size_t num_row, size_t num_col,
                          size_t col0, size_t B, size_t Ql, const double *fix, uint64_t *out, double2 *fftbuf,
                          hipStream_t s) {
    const size_t N = c.N, n = N / 2, M = 2 * N;
    if (B == 0) return;
    if (B > 65535) throw InvalidArgument("encode_columns_batch: at most 65535 plaintexts");
    if (num_x * num_row > n) throw InvalidArgument("batch_input: num_X * num_row exceeds the slot count");
    const int logn = ilog2(n), logRatio = ilog2(M) - logn - 2;
    const unsigned gx = std::min<unsigned>(grid_for(n / 2), 64u);
    k_columns_load_batch<<<dim3(std::min<unsigned>(grid_for(n), 128u), (unsigned)B), kBlock, 0, s>>>(
        X, (uint32_t)num_x, (uint32_t)num_row, (uint32_t)num_col, (uint32_t)col0, fftbuf, (uint32_t)n, logn);
    for (int iter = logn - 1; iter >= 0; iter--) {
        const int logPairs = logn - iter - 1;
        const int last = (size_t(1) << logPairs) == n / 2;
        k_ifft_stage_batch<<<dim3(gx, (unsigned)B), kBlock, 0, s>>>(fftbuf, c.enc_roots.as<double2>(),
                                                                     c.enc_group.as<uint32_t>(), (uint32_t)n, logn,
                                                                     logPairs, logRatio, (uint32_t)M, fix, last);
    }
    SHM_HIP(hipGetLastError());
    ntt_forward_encoded(c, fftbuf, out, B, Ql, s);
}

void encode_masked_batch(const Context &c, const double *w, const double *fix, const uint8_t *mask, size_t B,
                         size_t Ql, uint64_t *out, double2 *fftbuf, hipStream_t s) {
    const size_t N = c.N, n = N / 2, M = 2 * N;
    if (B == 0) return;
    if (B > 65535) throw InvalidArgument("encode_masked_batch: at most 65535 plaintexts");
    const int logn = ilog2(n), logRatio = ilog2(M) - logn - 2;
    const unsigned gx = std::min<unsigned>(grid_for(n / 2), 64u);
    k_mask_load_batch<<<dim3(std::min<unsigned>(grid_for(n), 128u), (unsigned)B), kBlock, 0, s>>>(w, mask, fftbuf,
                                                                                                 (uint32_t)n, logn);
    for (int iter = logn - 1; iter >= 0; iter--) {
        const int logPairs = logn - iter - 1;
        const int last = (size_t(1) << logPairs) == n / 2;
        k_ifft_stage_batch<<<dim3(gx, (unsigned)B), kBlock, 0, s>>>(fftbuf, c.enc_roots.as<double2>(),
                                                                     c.enc_group.as<uint32_t>(), (uint32_t)n, logn,
                                                                     logPairs, logRatio, (uint32_t)M, fix, last);
    }
    SHM_HIP(hipGetLastError());
    // decompose (round, residue per prime) fused into the NTT's column pass
    ntt_forward_encoded(c, fftbuf, out, B, Ql, s);
}

void encode_scalar(const Context &c, double value, size_t Ql, double scale, uint64_t *out, hipStream_t s) {
    if (!(scale > 0)) throw InvalidArgument("scale out of bounds");
    const double v = round(value * scale);
    k_fill_scalar<<<grid_for(Ql * c.N), kBlock, 0, s>>>(out, c.mods(), v, c.N, Ql * c.N);
    SHM_HIP(hipGetLastError());
}

void decode(const Context &c, const uint64_t *pt, size_t Ql, double scale, double *re, double *im, size_t nout,
            hipStream_t s) {
    const size_t N = c.N, n = N / 2, M = 2 * N;
    if (Ql > 64) throw InvalidArgument("decode supports at most 64 limbs");
    const int logn = ilog2(n), logRatio = ilog2(M) - logn - 2;
    // multi-precision constants for this level (host)
    const size_t W = Ql;
    std::vector<uint64_t> Q(W, 0), qhat(Ql * W, 0), qhinv(Ql), qhinv_s(Ql), half(W);
    auto mul_small = [&](uint64_t *a, uint64_t b) {
        hu128 carry = 0;
        for (size_t w = 0; w < W; w++) {
            hu128 t = (hu128)a[w] * b + carry;
            a[w] = (uint64_t)t;
            carry = t >> 64;
        }
    };
    Q[0] = 1;
    for (size_t i = 0; i < Ql; i++) mul_small(Q.data(), c.q[i]);
    for (size_t i = 0; i < Ql; i++) {
        uint64_t *h = qhat.data() + i * W;
        h[0] = 1;
        uint64_t r = 1;
        const uint64_t qi = c.q[i];
        for (size_t k = 0; k < Ql; k++)
            if (k != i) {
                mul_small(h, c.q[k]);
                r = (uint64_t)(((hu128)r * (c.q[k] % qi)) % qi);
            }
        // inverse by Fermat
        uint64_t e = qi - 2, b = r, inv = 1;
        while (e) {
            if (e & 1) inv = (uint64_t)(((hu128)inv * b) % qi);
            b = (uint64_t)(((hu128)b * b) % qi);
            e >>= 1;
        }
        qhinv[i] = inv;
        qhinv_s[i] = (uint64_t)(((hu128)inv << 64) / qi);
    }
    {
        std::vector<uint64_t> t = Q;
        hu128 carry = 1;
        for (size_t w = 0; w < W; w++) {
            hu128 sum = (hu128)t[w] + carry;
            t[w] = (uint64_t)sum;
            carry = sum >> 64;
        }
        for (size_t w = 0; w < W; w++) half[w] = (t[w] >> 1) | (w + 1 < W ? (t[w + 1] << 63) : 0);
    }
    const size_t cbytes = (W + Ql * W + 2 * Ql + W) * sizeof(uint64_t);
    Scratch consts(cbytes, s);
    uint64_t *dQ = consts.as<uint64_t>(), *dqhat = dQ + W, *dqhinv = dqhat + Ql * W, *dqhinv_s = dqhinv + Ql,
             *dhalf = dqhinv_s + Ql;
    SHM_HIP(hipMemcpyAsync(dQ, Q.data(), W * 8, hipMemcpyHostToDevice, s));
    SHM_HIP(hipMemcpyAsync(dqhat, qhat.data(), Ql * W * 8, hipMemcpyHostToDevice, s));
    SHM_HIP(hipMemcpyAsync(dqhinv, qhinv.data(), Ql * 8, hipMemcpyHostToDevice, s));
    SHM_HIP(hipMemcpyAsync(dqhinv_s, qhinv_s.data(), Ql * 8, hipMemcpyHostToDevice, s));
    SHM_HIP(hipMemcpyAsync(dhalf, half.data(), W * 8, hipMemcpyHostToDevice, s));

    Scratch x(Ql * N * sizeof(uint64_t), s);
    launch_copy_words(x.as<uint64_t>(), pt, Ql * N, s);
    ntt_inverse(c, x.as<uint64_t>(), 1, Ql, 0, s);
    Scratch buf(n * sizeof(double2), s);
    launch_fill_zero(buf.as<uint64_t>(), 2 * n, s);  // all-zero bits are +0.0
    k_compose<<<grid_for(2 * n), 64, 0, s>>>(x.as<uint64_t>(), buf.as<double2>(), c.mods(), dqhat, dqhinv, dqhinv_s,
                                            dQ, dhalf, Ql, N, (uint32_t)n, N / (2 * n), 1.0 / scale);
    for (int iter = 0; iter < logn; iter++) {
        const int logPairs = logn - iter - 1;
        k_fft_stage<<<grid_for(n / 2), kBlock, 0, s>>>(buf.as<double2>(), c.enc_roots.as<double2>(),
                                                       c.enc_group.as<uint32_t>(), (uint32_t)n, logn, logPairs,
                                                       logRatio, (uint32_t)M);
    }
    Scratch o(2 * n * sizeof(double), s);
    k_bitrev_store<<<grid_for(n), kBlock, 0, s>>>(buf.as<double2>(), o.as<double>(), o.as<double>() + n, (uint32_t)n,
                                                  logn);
    SHM_HIP(hipGetLastError());
    const size_t cnt = std::min(nout, n);
    SHM_HIP(hipMemcpyAsync(re, o.as<double>(), cnt * sizeof(double), hipMemcpyDeviceToHost, s));
    if (im) SHM_HIP(hipMemcpyAsync(im, o.as<double>() + n, cnt * sizeof(double), hipMemcpyDeviceToHost, s));
    SHM_HIP(hipStreamSynchronize(s));
}

}  // namespace shm
