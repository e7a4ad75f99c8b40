// Hybrid (dnum) key switching: ModUp -> key inner product -> ModDown.
//
// ModUp (rns_bconv.cu:530-634): c2 (Ql limbs, NTT) -> per digit g of alpha
// limbs: own limbs copied in NTT form; complement limbs of QlP get the
// approximate fast base conversion sum_i [x_i qhat_i^-1]_{q_i} [qhat_i]_p of the
// coefficient-form digit (alpha = 1: [x]_p), then an NTT.
// Inner product (eval_key_switch.cu:12-60): over QlP limbs,
// cx0 = sum_g d_g key_g.b, cx1 = sum_g d_g key_g.a with 128-bit accumulation.
// ModDown (rns_bconv.cu:776-828, ntt_moddown.cu:106-261): iNTT of the P limbs,
// delta = fastbconv(P -> Ql), NTT(delta), out = (cx - delta) P^-1 mod q.
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

// One launch covers every digit: out[g][k][j] for all g < beta, k < QlP.
__global__ void k_modup_bconv(const uint64_t *__restrict__ c2, const uint64_t *__restrict__ tcoef,
                              uint64_t *__restrict__ out, const DevMod *__restrict__ mods,
                              const uint64_t *__restrict__ hat_inv, const uint64_t *__restrict__ hat_inv_s,
                              const uint64_t *__restrict__ hat_mod, size_t Ql, size_t nQ, size_t alpha,
                              size_t beta, size_t N) {
    const size_t QlP = Ql + alpha;
    const size_t total = beta * QlP * N;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t g = t / (QlP * N), r = t % (QlP * N), k = r / N, j = r % N;
        const size_t start = g * alpha, size = min(alpha, Ql - start);
        if (k >= start && k < start + size) {
            out[t] = c2[k * N + j];
            continue;
        }
        const DevMod mp = mods[qlp_prime(k, Ql, nQ)];
        if (alpha == 1) {
            out[t] = reduce64(tcoef[start * N + j], mp);
        } else {
            Acc128 acc{0, 0};
            for (size_t i = 0; i < size; i++) {
                const size_t li = start + i;
                const uint64_t qi = mods[li].q;
                const uint64_t x = mul_shoup(tcoef[li * N + j], hat_inv[li], hat_inv_s[li], qi);
                mac128(acc, x, hat_mod[(g * alpha + i) * QlP + k]);
            }
            out[t] = reduce128(acc, mp);
        }
    }
}

__global__ void k_ks_inner(const uint64_t *__restrict__ d, const uint64_t *__restrict__ key,
                           uint64_t *__restrict__ cx, const DevMod *__restrict__ mods, size_t Ql, size_t nQ,
                           size_t QP, size_t alpha, size_t beta, size_t N) {
    const size_t QlP = Ql + alpha, total = QlP * N;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t k = t / N, j = t % N;
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

// delta[i][j] = fastbconv of the coefficient-form P limbs (cxP) to q_i.
__global__ void k_moddown_bconv(const uint64_t *__restrict__ cxP, uint64_t *__restrict__ delta,
                                const DevMod *__restrict__ mods, const uint64_t *__restrict__ phat_inv,
                                const uint64_t *__restrict__ phat_inv_s, const uint64_t *__restrict__ phat_mod,
                                size_t Ql, size_t nQ, size_t nP, size_t N) {
    const size_t total = Ql * N;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t i = t / N, j = t % N;
        const DevMod mq = mods[i];
        if (nP == 1) {
            delta[t] = reduce64(cxP[j], mq);
        } else {
            Acc128 acc{0, 0};
            for (size_t jj = 0; jj < nP; jj++) {
                const uint64_t p = mods[nQ + jj].q;
                const uint64_t x = mul_shoup(cxP[jj * N + j], phat_inv[jj], phat_inv_s[jj], p);
                mac128(acc, x, phat_mod[i * nP + jj]);
            }
            delta[t] = reduce128(acc, mq);
        }
    }
}

// out = (cx - delta) P^-1 mod q; then ct (+)= out.
__global__ void k_moddown_combine(const uint64_t *__restrict__ cx, const uint64_t *__restrict__ delta,
                                  uint64_t *__restrict__ ct, const DevMod *__restrict__ mods,
                                  const uint64_t *__restrict__ pinv, const uint64_t *__restrict__ pinv_s,
                                  size_t N, size_t total, int accumulate) {
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t i = t / N;
        const uint64_t q = mods[i].q;
        const uint64_t v = mul_shoup(sub_mod(cx[t], delta[t], q), pinv[i], pinv_s[i], q);
        ct[t] = accumulate ? add_mod(ct[t], v, q) : v;
    }
}

void moddown_into(const Context &c, uint64_t *cx, size_t Ql, uint64_t *ct, bool accumulate, uint64_t *delta,
                  hipStream_t s) {
    const size_t N = c.N, nP = c.nP, QlP = Ql + nP;
    ntt_inverse_qlp(c, cx, Ql, Ql, QlP, 1, QlP * N, s);
    k_moddown_bconv<<<grid_for(Ql * N), kBlock, 0, s>>>(cx + Ql * N, delta, c.mods(), c.phat_inv.as<uint64_t>(),
                                                       c.phat_inv_s.as<uint64_t>(), c.phat_mod.as<uint64_t>(), Ql,
                                                       c.nQ, nP, N);
    SHM_HIP(hipGetLastError());
    ntt_forward(c, delta, 1, Ql, 0, s);
    k_moddown_combine<<<grid_for(Ql * N), kBlock, 0, s>>>(cx, delta, ct, c.mods(), c.pinv.as<uint64_t>(),
                                                         c.pinv_s.as<uint64_t>(), N, Ql * N, accumulate ? 1 : 0);
    SHM_HIP(hipGetLastError());
}

}  // namespace

void keyswitch(const Context &c, const uint64_t *c2, const uint64_t *key, size_t Ql, uint64_t *ct0, uint64_t *ct1,
               bool add_to_ct1, hipStream_t s) {
    const size_t N = c.N, alpha = c.nP, QlP = Ql + alpha;
    const LevelConsts &L = c.level(Ql);
    const size_t beta = L.beta;
    if (beta > c.dnum) throw LogicError("key switching key has too few digits for this level");
    Scratch tcoef(Ql * N * sizeof(uint64_t), s);
    Scratch d(beta * QlP * N * sizeof(uint64_t), s);
    Scratch cx(2 * QlP * N * sizeof(uint64_t), s);
    Scratch delta(Ql * N * sizeof(uint64_t), s);
    SHM_HIP(hipMemcpyAsync(tcoef.as<uint64_t>(), c2, Ql * N * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
    ntt_inverse(c, tcoef.as<uint64_t>(), 1, Ql, 0, s);
    k_modup_bconv<<<grid_for(beta * QlP * N), kBlock, 0, s>>>(
        c2, tcoef.as<uint64_t>(), d.as<uint64_t>(), c.mods(), L.hat_inv.as<uint64_t>(), L.hat_inv_s.as<uint64_t>(),
        L.hat_mod.as<uint64_t>(), Ql, c.nQ, alpha, beta, N);
    SHM_HIP(hipGetLastError());
    for (size_t g = 0; g < beta; g++) {
        const size_t start = g * alpha, size = std::min(alpha, Ql - start);
        ntt_forward_qlp(c, d.as<uint64_t>() + g * QlP * N, Ql, 0, QlP, start, start + size, 1, QlP * N, s);
    }
    k_ks_inner<<<grid_for(QlP * N), kBlock, 0, s>>>(d.as<uint64_t>(), key, cx.as<uint64_t>(), c.mods(), Ql, c.nQ,
                                                    c.nQP(), alpha, beta, N);
    SHM_HIP(hipGetLastError());
    moddown_into(c, cx.as<uint64_t>(), Ql, ct0, true, delta.as<uint64_t>(), s);
    moddown_into(c, cx.as<uint64_t>() + QlP * N, Ql, ct1, add_to_ct1, delta.as<uint64_t>(), s);
}

}  // namespace shm
