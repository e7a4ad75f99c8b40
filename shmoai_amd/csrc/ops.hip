// Evaluator-level operations shared by the C ABI and the bootstrapping
// pipeline: key generation, Galois automorphisms with key switching, rescale,
// level dropping, and the helpers that keep levels and scales aligned.
#include <algorithm>
#include <cmath>
#include <vector>

#include "engine.h"

namespace shm {

uint64_t host_splitmix64(uint64_t key, uint64_t ctr) {
    uint64_t z = key + (ctr + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Exact-weight ternary secret by a partial Fisher-Yates shuffle of SplitMix64
// draws: position draw i from counter i, sign from counter N + i.
std::vector<int8_t> ternary_secret(size_t N, uint64_t seed, uint32_t hw) {
    if (hw == 0 || hw > N) throw InvalidArgument("hamming weight out of range");
    std::vector<uint32_t> perm(N);
    for (size_t i = 0; i < N; i++) perm[i] = (uint32_t)i;
    std::vector<int8_t> coeff(N, 0);
    for (size_t i = 0; i < hw; i++) {
        const size_t j = i + (size_t)(host_splitmix64(seed, i) % (uint64_t)(N - i));
        std::swap(perm[i], perm[j]);
        coeff[perm[i]] = (host_splitmix64(seed, N + i) & 1) ? -1 : 1;
    }
    return coeff;
}

// NTT-form secret over QP from signed coefficients (host int8 -> device).
void load_secret(Context &c, SecretKey &sk, const int8_t *coeffs) {
    const size_t N = c.N, QP = c.nQP();
    Scratch dc(N, nullptr);
    SHM_HIP(hipMemcpy(dc.as<void>(), coeffs, N, hipMemcpyHostToDevice));
    sk.s.alloc(QP * N * sizeof(uint64_t));
    launch_signed_to_rns(c, dc.as<int8_t>(), sk.s.as<uint64_t>(), c.nQ, true, nullptr);
    ntt_forward(c, sk.s.as<uint64_t>(), 1, QP, 0, nullptr);
    SHM_HIP(hipDeviceSynchronize());
}

// secretkey.cu:328-373: dnum digits of (-(a s + e) + P [newkey]_{digit}, a).
void gen_kswitch(Context &c, const SecretKey &sk, const uint64_t *newkey_dev, uint64_t seed, KSwitchKey &out) {
    const size_t N = c.N, QP = c.nQP(), dnum = c.dnum;
    out.ctx = &c;
    out.k.alloc(dnum * 2 * QP * N * sizeof(uint64_t));
    Scratch e(QP * N * sizeof(uint64_t), nullptr);
    for (size_t g = 0; g < dnum; g++) {
        uint64_t *b = out.k.as<uint64_t>() + g * 2 * QP * N, *a = b + QP * N;
        launch_sample_uniform(c, host_splitmix64(seed, 2 * g), a, c.nQ, true, nullptr);
        launch_sample_gaussian_rns(c, host_splitmix64(seed, 2 * g + 1), e.as<uint64_t>(), c.nQ, true, nullptr);
        ntt_forward(c, e.as<uint64_t>(), 1, QP, 0, nullptr);
        launch_ks_key_digit(c, sk.s.as<uint64_t>(), newkey_dev, b, a, e.as<uint64_t>(), g, nullptr);
    }
    SHM_HIP(hipDeviceSynchronize());
}

void apply_galois(Context &c, Ciphertext &ct, uint32_t elt, const GaloisKeys &gk, hipStream_t s, bool rounded) {
    if (ct.size != 2) throw InvalidArgument("encrypted size must be 2");
    auto it = gk.keys.find(elt);
    if (it == gk.keys.end()) throw InvalidArgument("Galois elt not present");
    const size_t Ql = ct.Ql(), N = c.N;
    const uint32_t *table = c.galois_table(elt, s);
    if (rounded) {
        // the exact centred ModUp commutes with the Galois permutation, so the
        // unpermuted c1 is raised and its digits are read through the permutation,
        // and c0 is added through it by ModDown's last pass (the hoisted form of one
        // step: the same words, no permuted copies)
        const size_t QlP = Ql + c.nP, QP = c.nQP(), beta = c.level(Ql).beta;
        Scratch d(beta * QlP * N * sizeof(uint64_t), s), cx(2 * QlP * N * sizeof(uint64_t), s);
        ks_modup(c, ct.poly(1), Ql, d.as<uint64_t>(), s, true);
        KeyRefs kr{};
        kr.base[0] = it->second->k.as<uint64_t>();
        kr.stride[0] = 2 * QP * N;
        kr.base[1] = kr.base[0] + QP * N;
        kr.stride[1] = 2 * QP * N;
        ks_inner_multi(c, d.as<uint64_t>(), Ql, beta, kr, 2, false, cx.as<uint64_t>(), s, table);
        Ciphertext o;
        o.ctx = &c;
        o.resize(2, ct.chain, s, false);  // the fused ModDown writes every word
        o.scale = ct.scale;
        o.ntt = true;
        PolyPack op{}, src{};
        op.p[0] = o.poly(0);
        op.p[1] = o.poly(1);
        src.p[0] = ct.poly(0);
        ks_moddown_pack(c, cx.as<uint64_t>(), 2, Ql, op, 1, s, true, &src, table);
        ct = std::move(o);
        return;
    }
    Scratch t0(Ql * N * sizeof(uint64_t), s), t1(Ql * N * sizeof(uint64_t), s);
    launch_permute(c, ct.poly(0), t0.as<uint64_t>(), table, Ql, s);
    launch_permute(c, ct.poly(1), t1.as<uint64_t>(), table, Ql, s);
    launch_copy_words(ct.poly(0), t0.as<uint64_t>(), Ql * N, s);
    keyswitch(c, t1.as<uint64_t>(), it->second->k.as<uint64_t>(), Ql, ct.poly(0), ct.poly(1), false, s, rounded);
}

// evaluate.cu:1650-1706 rotate_internal (NAF split, numth.h:17-34).
void rotate_internal(Context &c, Ciphertext &ct, int step, const GaloisKeys &gk, hipStream_t s, bool rounded) {
    if (step == 0) return;
    const uint32_t elt = c.elt_from_step(step);
    if (gk.keys.count(elt)) {
        apply_galois(c, ct, elt, gk, s, rounded);
        return;
    }
    std::vector<int> naf;
    {
        const bool sign = step < 0;
        int v = std::abs(step);
        for (int i = 0; v; i++) {
            const int zi = (v & 1) ? 2 - (v & 3) : 0;
            v = (v - zi) >> 1;
            if (zi) naf.push_back((sign ? -zi : zi) * (1 << i));
        }
    }
    if (naf.size() == 1) throw InvalidArgument("Galois key not present");
    for (int t : naf)
        if ((size_t)std::abs(t) != (c.N >> 1)) rotate_internal(c, ct, t, gk, s, rounded);
}

// Groups of ordinary ciphertexts held as one container of 2m polynomials
// (b_0, a_0, b_1, a_1, ...) at one level; every member is rotated by the same
// element with the same key, so one permute launch covers all polynomials and
// the key switch of all a_i is one keyswitch_batch (key words read once).
namespace {
uint64_t even_bits(size_t npolys) {
    uint64_t m = 0;
    for (size_t p = 0; p < npolys; p += 2) m |= uint64_t(1) << p;
    return m;
}
}  // namespace

void apply_galois_group(Context &c, Ciphertext &g, uint32_t elt, const GaloisKeys &gk, hipStream_t s, bool rounded) {
    if (g.size == 0 || g.size % 2 || g.size > (size_t)kPolyPack)
        throw InvalidArgument("a ciphertext group holds 1..32 pairs of polynomials");
    auto it = gk.keys.find(elt);
    if (it == gk.keys.end()) throw InvalidArgument("Galois elt not present");
    const size_t m = g.size / 2, Ql = g.Ql(), N = c.N;
    const uint32_t *table = c.galois_table(elt, s);
    if (rounded) {
        // as in apply_galois: the group's unpermuted a_i are raised once, the inner
        // product reads their digits through the permutation, and ModDown's last pass
        // adds every b_i through it (the hoisted group form of one step)
        const size_t QlP = Ql + c.nP, beta = c.level(Ql).beta;
        Scratch d(m * beta * QlP * N * sizeof(uint64_t), s), cx(g.size * QlP * N * sizeof(uint64_t), s);
        PolyPack a{};
        for (size_t k = 0; k < m; k++) a.p[k] = g.poly(2 * k + 1);
        ks_modup_batch(c, a, m, Ql, d.as<uint64_t>(), s, true);
        ks_inner_batch(c, d.as<uint64_t>(), m, Ql, it->second->k.as<uint64_t>(), cx.as<uint64_t>(), s, table, &a);
        Ciphertext o;
        o.ctx = &c;
        o.resize(g.size, g.chain, s, false);  // the fused ModDown writes every word
        o.scale = g.scale;
        o.ntt = true;
        PolyPack op{}, src{};
        for (size_t k = 0; k < g.size; k++) {
            op.p[k] = o.poly(k);
            src.p[k] = g.poly(k);
        }
        ks_moddown_pack(c, cx.as<uint64_t>(), g.size, Ql, op, even_bits(g.size), s, true, &src, table);
        g = std::move(o);
        return;
    }
    Scratch t(g.size * Ql * N * sizeof(uint64_t), s);
    launch_permute(c, g.data(), t.as<uint64_t>(), table, g.size * Ql, s);
    // b_i := sigma(b_i); a_i is overwritten by the switched part below
    launch_copy_words(g.data(), t.as<uint64_t>(), g.size * Ql * N, s);
    PolyPack c2{}, ct01{};
    for (size_t i = 0; i < m; i++) {
        c2.p[i] = t.as<uint64_t>() + (2 * i + 1) * Ql * N;
        ct01.p[2 * i] = g.poly(2 * i);
        ct01.p[2 * i + 1] = g.poly(2 * i + 1);
    }
    keyswitch_batch(c, c2, m, it->second->k.as<uint64_t>(), Ql, ct01, even_bits(g.size), s, rounded);
}

void rotate_internal_group(Context &c, Ciphertext &g, int step, const GaloisKeys &gk, hipStream_t s, bool rounded) {
    if (step == 0) return;
    const uint32_t elt = c.elt_from_step(step);
    if (gk.keys.count(elt)) {
        apply_galois_group(c, g, elt, gk, s, rounded);
        return;
    }
    std::vector<int> naf;
    {
        const bool sign = step < 0;
        int v = std::abs(step);
        for (int i = 0; v; i++) {
            const int zi = (v & 1) ? 2 - (v & 3) : 0;
            v = (v - zi) >> 1;
            if (zi) naf.push_back((sign ? -zi : zi) * (1 << i));
        }
    }
    if (naf.size() == 1) throw InvalidArgument("Galois key not present");
    for (int t : naf)
        if ((size_t)std::abs(t) != (c.N >> 1)) rotate_internal_group(c, g, t, gk, s, rounded);
}

void rotate_hoisted_group(Context &c, const Ciphertext &g, const std::vector<int> &steps, const GaloisKeys &gk,
                          const std::vector<Ciphertext *> &outs, hipStream_t s, bool rounded) {
    if (g.size == 0 || g.size % 2 || g.size > (size_t)kPolyPack)
        throw InvalidArgument("a ciphertext group holds 1..32 pairs of polynomials");
    if (steps.size() != outs.size()) throw InvalidArgument("one output per step");
    const size_t m = g.size / 2, Ql = g.Ql(), N = c.N, QlP = Ql + c.nP, beta = c.level(Ql).beta;
    Scratch d(m * beta * QlP * N * sizeof(uint64_t), s), cx(2 * m * QlP * N * sizeof(uint64_t), s);
    bool have_modup = false;
    PolyPack a{};
    for (size_t k = 0; k < m; k++) a.p[k] = g.poly(2 * k + 1);
    for (size_t i = 0; i < steps.size(); i++) {
        Ciphertext &o = *outs[i];
        const uint32_t elt = c.elt_from_step(steps[i]);
        auto it = gk.keys.find(elt);
        if (steps[i] == 0 || it == gk.keys.end()) {
            ct_clone(o, g, s);
            rotate_internal_group(c, o, steps[i], gk, s, rounded);
            continue;
        }
        if (!have_modup) {
            ks_modup_batch(c, a, m, Ql, d.as<uint64_t>(), s, rounded);
            have_modup = true;
        }
        const uint32_t *table = c.galois_table(elt, s);
        ks_inner_batch(c, d.as<uint64_t>(), m, Ql, it->second->k.as<uint64_t>(), cx.as<uint64_t>(), s, table, &a);
        o.ctx = &c;
        o.resize(g.size, g.chain, s, false);  // the fused ModDown writes every word
        o.scale = g.scale;
        o.ntt = true;
        // ModDown's last pass adds each b_i read through the Galois permutation (no
        // permuted copy in between) and writes each a_i
        PolyPack op{}, src{};
        for (size_t k = 0; k < g.size; k++) {
            op.p[k] = o.poly(k);
            src.p[k] = g.poly(k);
        }
        ks_moddown_pack(c, cx.as<uint64_t>(), g.size, Ql, op, even_bits(g.size), s, rounded, &src, table);
    }
}

void rotate_hoisted(Context &c, const Ciphertext &ct, const std::vector<int> &steps, const GaloisKeys &gk,
                    const std::vector<Ciphertext *> &outs, hipStream_t s, bool rounded) {
    if (ct.size != 2) throw InvalidArgument("encrypted size must be 2");
    if (steps.size() != outs.size()) throw InvalidArgument("one output per step");
    const size_t Ql = ct.Ql(), N = c.N, QlP = Ql + c.nP, QP = c.nQP(), beta = c.level(Ql).beta;
    Scratch d(beta * QlP * N * sizeof(uint64_t), s), cx(2 * QlP * N * sizeof(uint64_t), s);
    bool have_modup = false;
    for (size_t i = 0; i < steps.size(); i++) {
        Ciphertext &o = *outs[i];
        const uint32_t elt = c.elt_from_step(steps[i]);
        auto it = gk.keys.find(elt);
        if (steps[i] == 0 || it == gk.keys.end()) {
            ct_clone(o, ct, s);
            rotate_internal(c, o, steps[i], gk, s, rounded);
            continue;
        }
        if (!have_modup) {
            ks_modup(c, ct.poly(1), Ql, d.as<uint64_t>(), s, rounded);
            have_modup = true;
        }
        const uint32_t *table = c.galois_table(elt, s);
        KeyRefs kr{};
        kr.base[0] = it->second->k.as<uint64_t>();
        kr.stride[0] = 2 * QP * N;
        kr.base[1] = kr.base[0] + QP * N;
        kr.stride[1] = 2 * QP * N;
        ks_inner_multi(c, d.as<uint64_t>(), Ql, beta, kr, 2, false, cx.as<uint64_t>(), s, table);
        o.ctx = &c;
        o.resize(2, ct.chain, s, false);  // the fused ModDown writes every word
        o.scale = ct.scale;
        o.ntt = true;
        // c0 is added through the Galois permutation by ModDown's last pass
        PolyPack op{}, src{};
        op.p[0] = o.poly(0);
        op.p[1] = o.poly(1);
        src.p[0] = ct.poly(0);
        ks_moddown_pack(c, cx.as<uint64_t>(), 2, Ql, op, 1, s, rounded, &src, table);
    }
}

void hoisting_inplace(Context &c, Ciphertext &ct, const std::vector<int> &steps, const GaloisKeys &gk,
                      hipStream_t s) {
    if (ct.size != 2) throw InvalidArgument("ciphertext size must be 2");
    if (steps.empty()) throw InvalidArgument("steps is empty");
    const size_t Ql = ct.Ql(), N = c.N, QlP = Ql + c.nP, QP = c.nQP(), beta = c.level(Ql).beta;
    Scratch d(beta * QlP * N * sizeof(uint64_t), s), cx(2 * QlP * N * sizeof(uint64_t), s);
    Scratch c0(Ql * N * sizeof(uint64_t), s), t0(Ql * N * sizeof(uint64_t), s);
    ks_modup(c, ct.poly(1), Ql, d.as<uint64_t>(), s);
    for (size_t i = 0; i < steps.size(); i++) {
        const uint32_t elt = c.elt_from_step(steps[i]);
        auto it = gk.keys.find(elt);
        if (it == gk.keys.end()) throw InvalidArgument("Galois elt not present");
        const uint32_t *table = c.galois_table(elt, s);
        KeyRefs kr{};
        kr.base[0] = it->second->k.as<uint64_t>();
        kr.stride[0] = 2 * QP * N;
        kr.base[1] = kr.base[0] + QP * N;
        kr.stride[1] = 2 * QP * N;
        ks_inner_multi(c, d.as<uint64_t>(), Ql, beta, kr, 2, false, cx.as<uint64_t>(), s, table, i > 0);
        launch_permute(c, ct.poly(0), i ? t0.as<uint64_t>() : c0.as<uint64_t>(), table, Ql, s);
        if (i) launch_add(c, c0.as<uint64_t>(), t0.as<uint64_t>(), Ql, 1, s);
    }
    launch_copy_words(ct.poly(0), c0.as<uint64_t>(), Ql * N, s);
    ks_moddown(c, cx.as<uint64_t>(), 2, Ql, ct.data(), Ql * N, 1, s);
}

void rescale_inplace(Context &c, Ciphertext &ct, hipStream_t s, bool rounded) {
    const size_t Ql = ct.Ql();
    if (Ql < 2) throw LogicError("end of modulus switching chain reached");
    if (!ct.ntt) throw InvalidArgument("CKKS encrypted must be in NTT form");
    launch_rescale(c, ct.data(), ct.data(), Ql, ct.size, s, rounded);
    ct.chain += 1;
    ct.scale = ct.scale / (double)c.q[Ql - 1];
}

void mod_switch_drop(Context &c, Ciphertext &ct, size_t new_chain, hipStream_t s) {
    const size_t Ql = ct.Ql(), nl = c.ql_of_chain(new_chain);
    if (new_chain < ct.chain) throw InvalidArgument("cannot switch to a higher level");
    if (new_chain == ct.chain) return;
    // compact [size][Ql][N] -> [size][nl][N] through a staging copy (ranges overlap)
    if (ct.size > 1) {
        Scratch tmp(ct.size * nl * c.N * sizeof(uint64_t), s);
        launch_copy_limbs(c, ct.data(), Ql, tmp.as<uint64_t>(), nl, ct.size, s);
        launch_copy_words(ct.data(), tmp.as<uint64_t>(), ct.size * nl * c.N, s);
    }
    ct.chain = new_chain;
}


// ---- level / scale management ---------------------------------------------------
namespace {

constexpr int kBlockOps = 256;
inline unsigned grid_ops(size_t work) {
    size_t g = (work + kBlockOps - 1) / kBlockOps;
    if (g > (1u << 20)) g = 1u << 20;
    return (unsigned)(g ? g : 1);
}

// Bootstrapper.cu:3271-3284 kernel_modraise_inplace: limb 0 (coefficient form,
// prime q0) lifted to every prime; values above floor(q0/2) represent x - q0.
__global__ void k_modraise(const uint64_t *src0, uint64_t *dst, const DevMod *mods, size_t nl, size_t N) {
    const uint64_t q0 = mods[0].q;
    const size_t total = nl * N;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t j = t >> __builtin_ctzll(N), i = t & (N - 1);
        const DevMod m = mods[j];
        const uint64_t x = src0[i];
        uint64_t v = reduce64(x, m);
        if (x > (q0 >> 1)) v = sub_mod(v, reduce64(q0, m), m.q);
        dst[t] = v;
    }
}

// Multiply by the monomial X^(N/2): in NTT slot j the factor is
// psi^((2 brv(j) + 1) N/2) = iota (-1)^brv(j) with iota = psi^(N/2); brv(j) is odd
// exactly for j >= N/2. `negate` multiplies by X^(3N/2) = -X^(N/2).
__global__ void k_mono_half(uint64_t *a, const DevMod *mods, const uint64_t *iota, const uint64_t *iota_s, size_t Ql,
                            size_t N, size_t total, int negate) {
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t limb = (t >> __builtin_ctzll(N)) % Ql, j = t & (N - 1);
        const uint64_t q = mods[limb].q;
        uint64_t v = mul_shoup(a[t], iota[limb], iota_s[limb], q);
        const bool neg = (j >= N / 2) != (negate != 0);
        a[t] = neg ? neg_mod(v, q) : v;
    }
}

}  // namespace

void ct_clone(Ciphertext &dst, const Ciphertext &src, hipStream_t s) {
    dst.ctx = src.ctx;
    dst.resize(src.size, src.chain, s, false);  // the copy writes every word
    dst.scale = src.scale;
    dst.ntt = src.ntt;
    launch_copy_words(dst.data(), src.data(), src.size * src.Ql() * src.ctx->N, s);
}

void ct_add(Context &c, Ciphertext &a, const Ciphertext &b, hipStream_t s) {
    if (a.chain != b.chain || a.size != b.size) throw InvalidArgument("ct_add: level or size mismatch");
    launch_add(c, a.data(), b.data(), a.Ql(), a.size, s);
}

void ct_sub(Context &c, Ciphertext &a, const Ciphertext &b, hipStream_t s) {
    if (a.chain != b.chain || a.size != b.size) throw InvalidArgument("ct_sub: level or size mismatch");
    launch_sub(c, a.data(), b.data(), a.Ql(), a.size, s);
}

// a += constant v (encoded at a's exact scale).
void ct_add_const(Context &c, Ciphertext &a, double v, hipStream_t s) {
    // add_plain of encode_scalar(v, a.scale) without materialising the plaintext
    if (!(a.scale > 0)) throw InvalidArgument("scale out of bounds");
    launch_add_scalar(c, a.data(), round(v * a.scale), a.Ql(), s);
}

void ct_add_const_batch(Context &c, const std::vector<Ciphertext *> &a, double v, hipStream_t s) {
    for (size_t i0 = 0; i0 < a.size(); i0 += kPolyPack) {
        const size_t m = std::min<size_t>(kPolyPack, a.size() - i0);
        PolyPack pa{};
        for (size_t i = 0; i < m; i++) {
            const Ciphertext &x = *a[i0 + i];
            if (!(x.scale > 0)) throw InvalidArgument("scale out of bounds");
            if (x.chain != a[i0]->chain || x.scale != a[i0]->scale)
                throw InvalidArgument("batched constant: ciphertexts at different levels or scales");
            pa.p[i] = a[i0 + i]->data();  // polynomial 0
        }
        launch_pack_op(c, pa, PolyPack{}, m, round(v * a[i0]->scale), a[i0]->Ql(), 0, s);
    }
}

void ct_mul_const_batch(Context &c, const std::vector<Ciphertext *> &a, double v, double sigma, hipStream_t s) {
    if (!(sigma > 0)) throw InvalidArgument("scale out of bounds");
    for (size_t i0 = 0; i0 < a.size(); i0 += kPolyPack / 2) {
        const size_t m = std::min<size_t>(kPolyPack / 2, a.size() - i0);
        PolyPack pa{};
        for (size_t i = 0; i < m; i++) {
            Ciphertext &x = *a[i0 + i];
            if (x.size != 2 || x.chain != a[i0]->chain)
                throw InvalidArgument("batched constant: size-2 ciphertexts at one level");
            pa.p[2 * i] = x.poly(0);
            pa.p[2 * i + 1] = x.poly(1);
        }
        launch_pack_op(c, pa, PolyPack{}, 2 * m, round(v * sigma), a[i0]->Ql(), 1, s);
        for (size_t i = 0; i < m; i++) a[i0 + i]->scale *= sigma;
    }
}

void ct_add_batch(Context &c, const std::vector<Ciphertext *> &a, const std::vector<const Ciphertext *> &b,
                  hipStream_t s) {
    if (a.size() != b.size()) throw InvalidArgument("ct_add_batch: operand counts differ");
    for (size_t i0 = 0; i0 < a.size(); i0 += kPolyPack / 2) {
        const size_t m = std::min<size_t>(kPolyPack / 2, a.size() - i0);
        PolyPack pa{}, pb{};
        for (size_t i = 0; i < m; i++) {
            Ciphertext &x = *a[i0 + i];
            const Ciphertext &y = *b[i0 + i];
            if (x.chain != y.chain || x.size != 2 || y.size != 2 || x.chain != a[i0]->chain)
                throw InvalidArgument("ct_add_batch: size-2 ciphertexts at one level");
            pa.p[2 * i] = x.poly(0);
            pa.p[2 * i + 1] = x.poly(1);
            pb.p[2 * i] = y.poly(0);
            pb.p[2 * i + 1] = y.poly(1);
        }
        launch_pack_op(c, pa, pb, 2 * m, 0.0, a[i0]->Ql(), 2, s);
    }
}

// a <- a * [v encoded at scale sigma]; scale becomes a.scale * sigma.
void ct_mul_const(Context &c, Ciphertext &a, double v, double sigma, hipStream_t s) {
    // multiply_plain by encode_scalar(v, sigma) without materialising the plaintext
    if (!(sigma > 0)) throw InvalidArgument("scale out of bounds");
    launch_mul_scalar(c, a.data(), round(v * sigma), a.Ql(), a.size, s);
    a.scale *= sigma;
}

// a <- rescale(a * [v encoded at scale sigma]); scale becomes a.scale * sigma / q_last.
void ct_mul_const_rescale(Context &c, Ciphertext &a, double v, double sigma, hipStream_t s, bool rounded) {
    if (a.Ql() < 2) throw LogicError("end of modulus switching chain reached");
    ct_mul_const(c, a, v, sigma, s);
    rescale_inplace(c, a, s, rounded);
}

// rescale_inplace of several ciphertexts at one level as one pass: one gather of
// the last limbs, one iNTT launch and one fused NTT launch for all polynomials,
// each output written to a fresh buffer.
void rescale_batch(Context &c, const std::vector<Ciphertext *> &cts, hipStream_t s, bool rounded) {
    if (cts.empty()) return;
    const size_t Ql = cts[0]->Ql(), N = c.N;
    if (Ql < 2) throw LogicError("end of modulus switching chain reached");
    size_t np = 0;
    for (const Ciphertext *ct : cts) {
        if (ct->Ql() != Ql) throw InvalidArgument("rescale_batch: ciphertexts at different levels");
        if (!ct->ntt) throw InvalidArgument("CKKS encrypted must be in NTT form");
        np += ct->size;
    }
    if (np > (size_t)kPolyPack) throw InvalidArgument("rescale_batch: at most 64 polynomials");
    const size_t L1 = Ql - 1;
    PolyPack lastp{}, in{}, out{};
    std::vector<Ciphertext> fresh(cts.size());
    size_t p = 0;
    for (size_t i = 0; i < cts.size(); i++) {
        fresh[i].ctx = &c;
        fresh[i].resize(cts[i]->size, cts[i]->chain + 1, s, false);  // the fused NTT writes every word
        for (size_t k = 0; k < cts[i]->size; k++, p++) {
            in.p[p] = cts[i]->poly(k);
            lastp.p[p] = cts[i]->poly(k) + L1 * N;
            out.p[p] = fresh[i].poly(k);
        }
    }
    Scratch last(np * N * sizeof(uint64_t), s);
    Scratch delta(np * L1 * N * sizeof(uint64_t), s);
    ntt_inverse_gather(c, lastp, last.as<uint64_t>(), np, 1, L1, s);
    const LevelConsts &L = c.level(Ql);
    rescale_ntt_fused_pack(c, last.as<uint64_t>(), in, out, delta.as<uint64_t>(), np, L1,
                           L.qlast_inv.as<uint64_t>(), L.qlast_inv_s.as<uint64_t>(), s, rounded);
    for (size_t i = 0; i < cts.size(); i++) {
        fresh[i].scale = cts[i]->scale / (double)c.q[Ql - 1];
        fresh[i].ntt = true;
        *cts[i] = std::move(fresh[i]);
    }
}

// Bring a to (chain, scale): drop to one level above the target, then one
// constant multiply + rescale (the constant 1 encoded at scale * q_last / a.scale,
// so the result carries `scale` up to the rounding of that constant).
void ct_align(Context &c, Ciphertext &a, size_t chain, double scale, hipStream_t s, bool rounded) {
    if (a.chain > chain) throw LogicError("ct_align: ciphertext is already below the target level");
    if (a.chain == chain) {
        if (std::fabs(a.scale / scale - 1.0) > 1e-9) throw LogicError("ct_align: scales differ at equal levels");
        a.scale = scale;
        return;
    }
    if (a.chain + 1 < chain) mod_switch_drop(c, a, chain - 1, s);
    const double qlast = (double)c.q[a.Ql() - 1];
    ct_mul_const_rescale(c, a, 1.0, scale * qlast / a.scale, s, rounded);
    a.scale = scale;
}

// out = rescale(relinearize(a * b)) at the lower of the two operands' levels.
void ct_mul_relin_rescale(Context &c, Ciphertext &out, const Ciphertext &a, const Ciphertext &b,
                          const KSwitchKey &rk, hipStream_t s, bool rounded) {
    if (a.size != 2 || b.size != 2) throw InvalidArgument("multiply supports size-2 ciphertexts");
    // the rounded form merges the ModDown and the rescale; a group of one takes
    // exactly the operations every member of mul_relin_rescale_batch takes
    if (rounded) {
        mul_relin_rescale_batch(c, {&out}, {&a}, {&b}, rk, s, true);
        return;
    }
    // the operand at the higher level enters through its first Ql limbs (the
    // limbs a mod switch to the lower level keeps), with no copy
    const size_t chain = std::max(a.chain, b.chain);
    const Ciphertext *pa = &a, *pb = &b;
    const size_t Ql = c.ql_of_chain(chain);
    out.ctx = &c;
    out.resize(3, chain, s, false);  // tensor / square write d0, d1, d2
    if (pa == pb || (pa->data() == pb->data()))
        launch_square(c, pa->data(), out.data(), Ql, s, pa->Ql());
    else
        launch_tensor(c, pa->data(), pb->data(), out.data(), Ql, s, pa->Ql(), pb->Ql());
    out.scale = pa->scale * pb->scale;
    out.ntt = true;
    keyswitch(c, out.poly(2), rk.k.as<uint64_t>(), Ql, out.poly(0), out.poly(1), true, s, rounded);
    out.size = 2;
    rescale_inplace(c, out, s, rounded);
}

// ct_mul_relin_rescale over m independent products, with one key switch pass
// (keyswitch_batch) and one rescale pass (rescale_batch) for the group. The
// rounded form (inside bootstrapping) divides P d + ks(d2) by P q_last in one
// exact centred step (keyswitch_rescale_batch) instead of ModDown then rescale:
// one conversion and one forward NTT of Ql - 1 limbs per polynomial fewer.
void mul_relin_rescale_batch(Context &c, const std::vector<Ciphertext *> &out,
                             const std::vector<const Ciphertext *> &a, const std::vector<const Ciphertext *> &b,
                             const KSwitchKey &rk, hipStream_t s, bool rounded,
                             const std::vector<const Ciphertext *> *cheb) {
    const size_t m = out.size();
    if (m == 0) return;
    if (a.size() != m || b.size() != m || 2 * m > (size_t)kPolyPack)
        throw InvalidArgument("mul_relin_rescale_batch: 1..32 products");
    if (cheb && (!rounded || cheb->size() != m)) throw LogicError("Chebyshev step: rounded products, one term each");
    PolyPack c2{}, ct01{};
    size_t Ql = 0;
    for (size_t i = 0; i < m; i++) {
        if (a[i]->size != 2 || b[i]->size != 2) throw InvalidArgument("multiply supports size-2 ciphertexts");
        // an operand at a higher level enters through its first Ql limbs, uncopied
        const size_t chain = std::max(a[i]->chain, b[i]->chain);
        const Ciphertext *pa = a[i], *pb = b[i];
        const size_t ql = c.ql_of_chain(chain);
        if (i == 0) Ql = ql;
        if (ql != Ql) throw InvalidArgument("mul_relin_rescale_batch: products at different levels");
        Ciphertext &o = *out[i];
        o.ctx = &c;
        o.resize(3, chain, s, false);  // tensor / square write d0, d1, d2
        const bool sq = pa == pb || (pa->data() == pb->data());
        if (cheb) {
            // 2 a b - T: T times the integer round(s_a s_b / s_T) carries the product's
            // scale (relative rounding below 2^-52 of T's own scale)
            const Ciphertext *t = (*cheb)[i];
            uint64_t cval = 0;
            if (t) {
                if (t->size != 2 || t->Ql() < Ql) throw LogicError("Chebyshev step: term below the product level");
                const double cv = std::nearbyint(pa->scale * pb->scale / t->scale);
                if (!(cv >= 1.0 && cv < 9.2e18)) throw LogicError("Chebyshev step: scale ratio out of range");
                cval = (uint64_t)cv;
            }
            launch_tensor_cheb(c, pa->data(), sq ? pa->data() : pb->data(), t ? t->data() : nullptr, cval, o.data(),
                               Ql, s, pa->Ql(), sq ? pa->Ql() : pb->Ql(), t ? t->Ql() : 0);
        } else if (sq) {
            launch_square(c, pa->data(), o.data(), Ql, s, pa->Ql());
        } else {
            launch_tensor(c, pa->data(), pb->data(), o.data(), Ql, s, pa->Ql(), pb->Ql());
        }
        o.scale = pa->scale * pb->scale;
        o.ntt = true;
        c2.p[i] = o.poly(2);
        ct01.p[2 * i] = o.poly(0);
        ct01.p[2 * i + 1] = o.poly(1);
    }
    if (rounded) {
        if (Ql < 2) throw LogicError("end of modulus switching chain reached");
        std::vector<Ciphertext> fresh(m);
        PolyPack outp{};
        for (size_t i = 0; i < m; i++) {
            fresh[i].ctx = &c;
            fresh[i].resize(2, out[i]->chain + 1, s, false);  // the fused NTT writes every word
            outp.p[2 * i] = fresh[i].poly(0);
            outp.p[2 * i + 1] = fresh[i].poly(1);
        }
        keyswitch_rescale_batch(c, c2, ct01, m, rk.k.as<uint64_t>(), Ql, outp, s);
        const double qlast = (double)c.q[Ql - 1];
        for (size_t i = 0; i < m; i++) {
            fresh[i].scale = out[i]->scale / qlast;
            fresh[i].ntt = true;
            *out[i] = std::move(fresh[i]);
        }
        return;
    }
    const uint64_t all = 2 * m == 64 ? ~uint64_t(0) : ((uint64_t(1) << (2 * m)) - 1);
    keyswitch_batch(c, c2, m, rk.k.as<uint64_t>(), Ql, ct01, all, s, rounded);  // relinearise: add to both
    for (Ciphertext *o : out) o->size = 2;
    rescale_batch(c, out, s, rounded);
}

namespace {
// out[k] = rotate_vector(in[k], step) for every k: the permutations, then one
// batched key switch per 32 ciphertexts (the key read once per batch). Each
// result equals rotate_internal's limb for limb; a step without its own key
// takes rotate_internal's NAF split one ciphertext at a time.
void rotate_many(Context &c, const std::vector<const Ciphertext *> &in, const std::vector<Ciphertext *> &out, int step,
                 const GaloisKeys &gk, hipStream_t s) {
    const uint32_t elt = c.elt_from_step(step);
    auto it = gk.keys.find(elt);
    if (step == 0 || it == gk.keys.end()) {
        for (size_t k = 0; k < in.size(); k++) {
            ct_clone(*out[k], *in[k], s);
            rotate_internal(c, *out[k], step, gk, s);
        }
        return;
    }
    const uint32_t *table = c.galois_table(elt, s);
    for (size_t k0 = 0; k0 < in.size(); k0 += 32) {
        const size_t m = std::min<size_t>(32, in.size() - k0);
        const size_t Ql = in[k0]->Ql();
        PolyPack c2{}, ct01{};
        for (size_t k = 0; k < m; k++) {
            const Ciphertext &x = *in[k0 + k];
            Ciphertext &o = *out[k0 + k];
            if (x.size != 2 || x.Ql() != Ql) throw InvalidArgument("rotate_many: size-2 ciphertexts at one level");
            o.ctx = &c;
            o.resize(2, x.chain, s, false);  // both polynomials are written by the permutation
            o.scale = x.scale;
            o.ntt = true;
            launch_permute(c, x.poly(0), o.poly(0), table, Ql, s);
            launch_permute(c, x.poly(1), o.poly(1), table, Ql, s);
            // the permuted c1 is the key switch input; ModDown overwrites it with
            // the new c1 after the inner products have read it
            c2.p[k] = o.poly(1);
            ct01.p[2 * k] = o.poly(0);
            ct01.p[2 * k + 1] = o.poly(1);
        }
        const uint64_t add_c0 = 0x5555555555555555ull & ((m == 32) ? ~uint64_t(0) : (uint64_t(1) << (2 * m)) - 1);
        keyswitch_batch(c, c2, m, it->second->k.as<uint64_t>(), Ql, ct01, add_c0, s, false);
    }
}

// relinearize of 3-polynomial ciphertexts at one level (Ql limbs), batched when
// the level's digit count allows it (same result either way)
void relin_many(Context &c, const std::vector<Ciphertext *> &cts, const KSwitchKey &rk, size_t Ql, hipStream_t s) {
    const size_t m = cts.size();
    PolyPack c2{}, ct01{};
    for (size_t i = 0; i < m; i++) {
        c2.p[i] = cts[i]->poly(2);
        ct01.p[2 * i] = cts[i]->poly(0);
        ct01.p[2 * i + 1] = cts[i]->poly(1);
    }
    const uint64_t all = 2 * m == 64 ? ~uint64_t(0) : (uint64_t(1) << (2 * m)) - 1;
    keyswitch_batch(c, c2, m, rk.k.as<uint64_t>(), Ql, ct01, all, s, false);  // relinearise: add to both
    for (Ciphertext *o : cts) o->size = 2;
}
}  // namespace

// Ct_ct_matrix_mul.cuh:10-181 ct_ct_matrix_mul_colpacking. Output i is
// rescale(relinearize(sum_j X[j] * rot(W[j], i * num_batch))): the reference
// multiplies (tensor, no relinearisation), adds the col_X products and then
// relinearises and rescales once, and so does this. Each rotation is the
// reference's rotate_vector (permutation, then key switch of the permuted c1),
// so the result is the same limb for limb; the col_X rotations of one output
// share a step and go through batched key switches. Every product of an output
// accumulates into its 3-polynomial ciphertext in the tensor kernel, and the
// outputs are relinearised and rescaled in batched passes of up to 16. Steps of
// the slot count or more are refused, as get_elt_from_step does (galois.cu:64-69).
void ctct_matmul_colpacking(Context &c, const std::vector<const Ciphertext *> &X,
                            const std::vector<const Ciphertext *> &W, int num_batch, const GaloisKeys &gk,
                            const KSwitchKey &rk, const std::vector<Ciphertext *> &out, hipStream_t s) {
    const size_t colX = X.size(), rowX = out.size();
    if (colX == 0 || W.size() != colX) throw InvalidArgument("ct_ct_matrix_mul: X and W need col_X ciphertexts each");
    if (rowX == 0) throw InvalidArgument("ct_ct_matrix_mul: row_X must be positive");
    if (num_batch < 0) throw InvalidArgument("ct_ct_matrix_mul: num_batch must be non-negative");
    const size_t chain = X[0]->chain;
    for (size_t j = 0; j < colX; j++) {
        for (const Ciphertext *x : {X[j], W[j]}) {
            if (!x || x->size != 2 || !x->ntt) throw InvalidArgument("ct_ct_matrix_mul: size-2 NTT-form ciphertexts");
            if (x->chain != chain) throw InvalidArgument("ct_ct_matrix_mul: operands at different levels");
        }
        if (X[j]->scale != X[0]->scale || W[j]->scale != W[0]->scale)
            throw InvalidArgument("ct_ct_matrix_mul: operands at different scales");
    }
    for (size_t i = 0; i < rowX; i++) {
        if (!out[i]) throw InvalidArgument("ct_ct_matrix_mul: null output");
        for (size_t j = 0; j < colX; j++)
            if (out[i] == X[j] || out[i] == W[j]) throw InvalidArgument("ct_ct_matrix_mul: output aliases an input");
        for (size_t k = 0; k < i; k++)
            if (out[k] == out[i]) throw InvalidArgument("ct_ct_matrix_mul: an output appears twice");
    }
    const size_t Ql = c.ql_of_chain(chain), slots = c.N / 2;
    if (Ql < 2) throw LogicError("end of modulus switching chain reached");
    if ((rowX - 1) * (size_t)num_batch >= slots) throw InvalidArgument("step count too large");
    constexpr size_t kChunk = 16;
    std::vector<Ciphertext> rot(colX);
    std::vector<Ciphertext *> rotp(colX);
    for (size_t j = 0; j < colX; j++) rotp[j] = &rot[j];
    for (size_t i0 = 0; i0 < rowX; i0 += kChunk) {
        const size_t i1 = std::min(rowX, i0 + kChunk);
        for (size_t i = i0; i < i1; i++) {
            Ciphertext &o = *out[i];
            o.ctx = &c;
            o.resize(3, chain, s, false);  // the first product writes every word
            o.ntt = true;
            o.scale = X[0]->scale * W[0]->scale;
        }
        for (size_t i = i0; i < i1; i++) {
            const size_t step = i * (size_t)num_batch;
            if (step != 0) rotate_many(c, W, rotp, (int)step, gk, s);  // every W[j] by the same step
            for (size_t j = 0; j < colX; j++) {
                const Ciphertext *w = step != 0 ? &rot[j] : W[j];
                launch_tensor(c, X[j]->data(), w->data(), out[i]->data(), Ql, s, X[j]->Ql(), w->Ql(), j > 0);
            }
        }
        std::vector<Ciphertext *> chunk(out.begin() + i0, out.begin() + i1);
        relin_many(c, chunk, rk, Ql, s);
        rescale_batch(c, chunk, s, false);
    }
}

// Ct_ct_matrix_mul.cuh:182-609 ct_ct_matrix_mul_diagpacking (baby-step giant-step):
// g = ceil(sqrt(col_X)), b = ceil(col_X / g). X[j g + k] is pre-rotated by
// (col_X - j g) num_batch (j > 0); for each output i the babies are rot(W[i], k
// num_batch), k < g; giant j sums the products of baby k with X[j g + k] (tensor,
// accumulated), is relinearised and rescaled and labelled with the scale of X, as
// the reference does; output i = sum_j rot(giant_j, j g num_batch). Rotations are
// rotate_vector's, so the result is the reference's limb for limb. The giants of
// one output are relinearised and rescaled in one batched pass.
void ctct_matmul_diagpacking(Context &c, const std::vector<const Ciphertext *> &X,
                             const std::vector<const Ciphertext *> &W, int num_batch, const GaloisKeys &gk,
                             const KSwitchKey &rk, const std::vector<Ciphertext *> &out, hipStream_t s) {
    const size_t colX = X.size(), colW = W.size();
    if (colX == 0 || colW == 0) throw InvalidArgument("ct_ct_matrix_mul_diag: empty operands");
    if (out.size() != colW) throw InvalidArgument("ct_ct_matrix_mul_diag: one output per W ciphertext");
    if (num_batch < 0) throw InvalidArgument("ct_ct_matrix_mul: num_batch must be non-negative");
    const size_t chain = X[0]->chain;
    auto check_ct = [&](const Ciphertext *x) {
        if (!x || x->size != 2 || !x->ntt) throw InvalidArgument("ct_ct_matrix_mul: size-2 NTT-form ciphertexts");
        if (x->chain != chain) throw InvalidArgument("ct_ct_matrix_mul: operands at different levels");
    };
    for (const Ciphertext *x : X) {
        check_ct(x);
        if (x->scale != X[0]->scale) throw InvalidArgument("ct_ct_matrix_mul: operands at different scales");
    }
    for (const Ciphertext *w : W) {
        check_ct(w);
        if (w->scale != W[0]->scale) throw InvalidArgument("ct_ct_matrix_mul: operands at different scales");
    }
    for (size_t i = 0; i < colW; i++) {
        if (!out[i]) throw InvalidArgument("ct_ct_matrix_mul: null output");
        for (const Ciphertext *x : X)
            if (out[i] == x) throw InvalidArgument("ct_ct_matrix_mul: output aliases an input");
        for (const Ciphertext *w : W)
            if (out[i] == w) throw InvalidArgument("ct_ct_matrix_mul: output aliases an input");
        for (size_t k = 0; k < i; k++)
            if (out[k] == out[i]) throw InvalidArgument("ct_ct_matrix_mul: an output appears twice");
    }
    const size_t Ql = c.ql_of_chain(chain), slots = c.N / 2, nb = (size_t)num_batch;
    if (Ql < 2) throw LogicError("end of modulus switching chain reached");
    size_t g = (size_t)std::sqrt((double)colX);
    if (g * g < colX) g++;
    size_t b = colX / g;
    if (b * g < colX) b++;
    if (b > 32) throw InvalidArgument("ct_ct_matrix_mul_diag: col_X above 1024");
    // every rotation step stays below the slot count (get_elt_from_step, galois.cu:64-69)
    if ((b > 1 && (colX - g) * nb >= slots) || (g - 1) * nb >= slots || (b - 1) * g * nb >= slots)
        throw InvalidArgument("step count too large");
    auto rot_step = [&](size_t step) { return (int)step; };
    // pre-rotated X (giant 0 uses X as it is)
    std::vector<Ciphertext> rx(colX);
    std::vector<const Ciphertext *> xr(colX);
    for (size_t j = 0; j < b; j++) {
        const size_t lo = j * g, hi = std::min(colX, lo + g);
        const int st = rot_step((colX - j * g) * nb);
        if (j == 0 || st == 0) {
            for (size_t idx = lo; idx < hi; idx++) xr[idx] = X[idx];
            continue;
        }
        // the group of giant j shares one step: one batched rotation
        std::vector<const Ciphertext *> src(X.begin() + lo, X.begin() + hi);
        std::vector<Ciphertext *> dst;
        for (size_t idx = lo; idx < hi; idx++) {
            dst.push_back(&rx[idx]);
            xr[idx] = &rx[idx];
        }
        rotate_many(c, src, dst, st, gk, s);
    }
    std::vector<Ciphertext> baby(g), giant(b);
    std::vector<const Ciphertext *> bp(g);
    for (size_t i = 0; i < colW; i++) {
        bp[0] = W[i];
        for (size_t k = 1; k < g; k++) {
            const int st = rot_step(k * nb);
            if (st == 0) {
                bp[k] = W[i];
                continue;
            }
            ct_clone(baby[k], *W[i], s);
            rotate_internal(c, baby[k], st, gk, s);
            bp[k] = &baby[k];
        }
        std::vector<Ciphertext *> gs(b);
        for (size_t j = 0; j < b; j++) {
            Ciphertext &o = giant[j];
            o.ctx = &c;
            o.resize(3, chain, s, false);  // the first product writes every word
            o.ntt = true;
            o.scale = W[0]->scale * X[0]->scale;
            for (size_t k = 0; k < g && j * g + k < colX; k++)
                launch_tensor(c, bp[k]->data(), xr[j * g + k]->data(), o.data(), Ql, s, bp[k]->Ql(),
                              xr[j * g + k]->Ql(), k > 0);
            gs[j] = &o;
        }
        relin_many(c, gs, rk, Ql, s);
        rescale_batch(c, gs, s, false);
        for (Ciphertext *o : gs) o->scale = X[0]->scale;
        ct_clone(*out[i], giant[0], s);
        for (size_t j = 1; j < b; j++) {
            rotate_internal(c, giant[j], rot_step(j * g * nb), gk, s);
            ct_add(c, *out[i], giant[j], s);
        }
    }
}

void ct_mul_monomial_half(Context &c, Ciphertext &a, bool negate, hipStream_t s) {
    const size_t total = a.size * a.Ql() * c.N;
    k_mono_half<<<grid_ops(total), kBlockOps, 0, s>>>(a.data(), c.mods(), c.iota.as<uint64_t>(),
                                                      c.iota_s.as<uint64_t>(), a.Ql(), c.N, total, negate ? 1 : 0);
    SHM_HIP(hipGetLastError());
}

// Bootstrapper.cu:3287-3356 modraise_inplace: bottom-level (1 prime) ciphertext
// lifted to all data primes; the result is in NTT form at chain index 1. Every
// polynomial is lifted, so a shared-a batch (SharedABatchOps.cu modraise) goes
// through the same path.
void ct_modraise(Context &c, Ciphertext &a, hipStream_t s) {
    if (a.size < 2) throw InvalidArgument("Ciphertexts of size 2 are supported only!");
    if (a.Ql() != 1) throw InvalidArgument("Ciphertexts in the lowest level are supported only!");
    const size_t N = c.N, nl = c.nQ, np = a.size;
    Scratch low(np * N * sizeof(uint64_t), s);
    launch_copy_words(low.as<uint64_t>(), a.data(), np * N, s);
    ntt_inverse(c, low.as<uint64_t>(), np, 1, 0, s);
    Ciphertext out;
    out.ctx = &c;
    out.resize(np, 1, s, false);  // k_modraise writes every limb
    for (size_t p = 0; p < np; p++)
        k_modraise<<<grid_ops(nl * N), kBlockOps, 0, s>>>(low.as<uint64_t>() + p * N, out.poly(p), c.mods(), nl, N);
    SHM_HIP(hipGetLastError());
    ntt_forward(c, out.data(), np, nl, 0, s);
    out.scale = a.scale;
    a = std::move(out);
}

}  // namespace shm
