// Shared-a ciphertext batches (n members that share one a-part, each under its
// own auxiliary secret s'_i) and the operations of shared-a S2C-first
// bootstrapping:
//   FmtSwitch forward (FmtSwitch.cu:276-394): n ordinary ciphertexts ->
//     a' = ModDown(sum_j ext(a_j) a_fmt[j]), b'_i = b_i + ModDown(sum_j ext(a_j) B[j][i])
//     with ext the exact Ql -> QlP extension of a_j;
//   FmtSwitch backward (FmtSwitch.cu:396-429): member i -> (b_i, 0) + KS_{s'_i -> s}(a);
//   Sh-Rot (ShRotate.cu:201-257): sigma of every polynomial, one ModUp of sigma(a),
//     n + 1 inner products with the shared Galois key, n + 1 ModDowns.
// A batch is stored as one [n+1][Ql][N] block: polynomials 0..n-1 are the b_i,
// polynomial n is a. Elementwise batch operations (SharedABatchOps.cu) are then
// single launches over all members.
#include <cmath>

#include "arith.h"
#include "boot.h"

namespace shm {
namespace boot {

namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(size_t work) {
    size_t g = (work + kBlock - 1) / kBlock;
    if (g > (1u << 20)) g = 1u << 20;
    return (unsigned)(g ? g : 1);
}

uint64_t h_mulmod(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)((unsigned __int128)a * b % q); }
uint64_t h_pow(uint64_t b, uint64_t e, uint64_t q) {
    uint64_t r = 1 % q;
    b %= q;
    while (e) {
        if (e & 1) r = h_mulmod(r, b, q);
        b = h_mulmod(b, b, q);
        e >>= 1;
    }
    return r;
}
uint64_t h_shoup(uint64_t w, uint64_t q) { return (uint64_t)(((unsigned __int128)w << 64) / q); }

// FmtSwitch.cu:82-108 extend_Ql_to_P_kernel: exact base extension of coefficient
// residues x_g (g < Ql) to the special primes. y_g = [x_g (Ql/q_g)^-1]_{q_g},
// x = sum_g y_g (Ql/q_g) - v Ql with v = floor(sum_g y_g / q_g), the floor taken
// in double with a guard that rounds values within 1e-9 below an integer up.
__global__ void k_extend_exact(const uint64_t *__restrict__ coef, uint64_t *__restrict__ outP,
                               const DevMod *__restrict__ mods, const uint64_t *__restrict__ inv_g,
                               const uint64_t *__restrict__ inv_g_s, const uint64_t *__restrict__ cg,
                               const uint64_t *__restrict__ qmodp, size_t Ql, size_t nQ, size_t nP, size_t N,
                               size_t npolys, size_t in_stride, size_t out_stride) {
    const size_t total = npolys * N;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t p = t >> __builtin_ctzll(N), j = t & (N - 1);
        const uint64_t *x = coef + p * in_stride;
        double v = 0.0;
        for (size_t g = 0; g < Ql; g++) {
            const uint64_t q = mods[g].q;
            const uint64_t y = mul_shoup(x[g * N + j], inv_g[g], inv_g_s[g], q);
            v = __dadd_rn(v, __ddiv_rn((double)y, (double)q));
        }
        uint64_t alpha = (uint64_t)v;
        if (__dsub_rn(v, (double)alpha) > 1.0 - 1e-9) alpha += 1;
        for (size_t k = 0; k < nP; k++) {
            const DevMod mp = mods[nQ + k];
            Acc128 acc{0, 0};
            for (size_t g = 0; g < Ql; g++) {
                const uint64_t q = mods[g].q;
                const uint64_t y = mul_shoup(x[g * N + j], inv_g[g], inv_g_s[g], q);
                mac128(acc, y, cg[k * Ql + g]);
            }
            const uint64_t am = reduce128(acc, mp);
            const uint64_t corr = mul_mod(reduce64(alpha, mp), qmodp[k], mp);
            outP[p * out_stride + k * N + j] = sub_mod(am, corr, mp.q);
        }
    }
}

// FmtSwitch.cu:160-270 gen_format_key entry over QlP (local limb order):
// B = -(a * s_aux + e) (+ P * s_main on the Ql limbs when diag).
__global__ void k_fmt_key_entry(const uint64_t *__restrict__ a, const uint64_t *__restrict__ s_aux,
                                const uint64_t *__restrict__ e, const uint64_t *__restrict__ s_main,
                                uint64_t *__restrict__ B, const DevMod *__restrict__ mods,
                                const uint64_t *__restrict__ pmod, const uint64_t *__restrict__ pmod_s, size_t Ql,
                                size_t nQ, size_t nP, size_t N, int diag) {
    const size_t total = (Ql + nP) * N;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t k = t >> __builtin_ctzll(N), j = t & (N - 1);
        const size_t pi = k < Ql ? k : nQ + (k - Ql);
        const DevMod m = mods[pi];
        const uint64_t as = mul_mod(a[t], s_aux[pi * N + j], m);
        uint64_t v = neg_mod(add_mod(as, e[t], m.q), m.q);
        if (diag && k < Ql) v = add_mod(v, mul_shoup(s_main[pi * N + j], pmod[pi], pmod_s[pi], m.q), m.q);
        B[t] = v;
    }
}

size_t qlp_bytes(const Context &c, size_t Ql) { return (Ql + c.nP) * c.N * sizeof(uint64_t); }

}  // namespace

// ---- key material ----------------------------------------------------------------------

std::unique_ptr<SharedAKeys> make_shared_keys(Context &c, const SecretKey &sk, size_t n, uint32_t hw,
                                              uint64_t seed, size_t fmt_chain, hipStream_t s) {
    if (n == 0) throw InvalidArgument("SharedAKeyGen: batch size must be positive");
    if (n > 16) throw InvalidArgument("shared-a batches hold at most 16 members");
    if (fmt_chain == 0 || fmt_chain > c.nQ) throw InvalidArgument("SharedAKeyGen::gen_format_key: invalid chain index");
    const size_t N = c.N, nP = c.nP, QP = c.nQP(), Ql = c.ql_of_chain(fmt_chain), QlP = Ql + nP;
    auto K = std::make_unique<SharedAKeys>();
    K->ctx = &c;
    K->n = n;
    K->fmt_chain = fmt_chain;
    K->hw = hw;
    K->seed = seed;
    // auxiliary secrets s'_i (SharedAKeyGen ctor, FmtSwitch.cu:154-163)
    K->aux.resize(n);
    K->aux_coeffs.resize(n);
    for (size_t i = 0; i < n; i++) {
        K->aux_coeffs[i] = ternary_secret(N, host_splitmix64(seed, i), hw);
        K->aux[i].ctx = &c;
        load_secret(c, K->aux[i], K->aux_coeffs[i].data());
    }
    // exact-extension constants of the format level
    {
        std::vector<uint64_t> inv(Ql), invs(Ql), cg(nP * Ql), qmp(nP);
        for (size_t g = 0; g < Ql; g++) {
            uint64_t hat = 1;
            for (size_t h = 0; h < Ql; h++)
                if (h != g) hat = h_mulmod(hat, c.q[h] % c.q[g], c.q[g]);
            inv[g] = h_pow(hat, c.q[g] - 2, c.q[g]);
            invs[g] = h_shoup(inv[g], c.q[g]);
        }
        for (size_t k = 0; k < nP; k++) {
            const uint64_t p = c.q[c.nQ + k];
            uint64_t all = 1;
            for (size_t g = 0; g < Ql; g++) all = h_mulmod(all, c.q[g] % p, p);
            qmp[k] = all;
            for (size_t g = 0; g < Ql; g++) {
                uint64_t hat = 1;
                for (size_t h = 0; h < Ql; h++)
                    if (h != g) hat = h_mulmod(hat, c.q[h] % p, p);
                cg[k * Ql + g] = hat;
            }
        }
        auto up = [](DevBuf &b, const std::vector<uint64_t> &v) {
            b.alloc(v.size() * sizeof(uint64_t));
            SHM_HIP(hipMemcpy(b.as<void>(), v.data(), v.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
        };
        up(K->ext_inv, inv);
        up(K->ext_inv_s, invs);
        up(K->ext_cg, cg);
        up(K->ext_qmodp, qmp);
    }
    // format key (FmtSwitch.cu:164-262): a_fmt[j] uniform over QlP; B[j][i]
    const uint64_t fseed = host_splitmix64(seed, n);
    K->a_fmt.alloc(n * qlp_bytes(c, Ql));
    K->B_fmt.alloc(n * n * qlp_bytes(c, Ql));
    Scratch e(qlp_bytes(c, Ql), s);
    for (size_t j = 0; j < n; j++) {
        uint64_t *aj = K->a_fmt.as<uint64_t>() + j * QlP * N;
        launch_sample_uniform(c, host_splitmix64(fseed, j), aj, Ql, true, s);
        for (size_t i = 0; i < n; i++) {
            launch_sample_gaussian_rns(c, host_splitmix64(fseed, n + j * n + i), e.as<uint64_t>(), Ql, true, s);
            ntt_forward(c, e.as<uint64_t>(), 1, Ql, 0, s);
            ntt_forward(c, e.as<uint64_t>() + Ql * N, 1, nP, c.nQ, s);
            k_fmt_key_entry<<<grid_for(QlP * N), kBlock, 0, s>>>(
                aj, K->aux[i].s.as<uint64_t>(), e.as<uint64_t>(), sk.s.as<uint64_t>(),
                K->B_fmt.as<uint64_t>() + (j * n + i) * QlP * N, c.mods(), c.pmod.as<uint64_t>(),
                c.pmod_s.as<uint64_t>(), Ql, c.nQ, nP, N, i == j ? 1 : 0);
            SHM_HIP(hipGetLastError());
        }
    }
    // backward keys s'_i -> s (FmtSwitch.cu:264-268)
    K->backward.resize(n);
    for (size_t i = 0; i < n; i++) {
        K->backward[i] = std::make_unique<KSwitchKey>();
        gen_kswitch(c, sk, K->aux[i].s.as<uint64_t>(), host_splitmix64(seed, 2 * n + 2 + i), *K->backward[i]);
    }
    SHM_HIP(hipStreamSynchronize(s));
    (void)QP;
    return K;
}

// ShRotate.cu:54-162 gen_shared_galois_key: A_g uniform over QP (shared), and per
// member B_{g,i} = -(A_g s'_i + e) + P [sigma(s'_i)]_{digit g}.
void add_shared_galois_key(SharedAKeys &K, int step, uint64_t seed, hipStream_t s) {
    Context &c = *K.ctx;
    const uint32_t elt = c.elt_from_step(step);
    if (K.galois.count(elt)) return;
    const size_t N = c.N, QP = c.nQP(), dnum = c.dnum, n = K.n;
    auto key = std::make_unique<SharedGaloisKey>();
    key->elt = elt;
    key->A.alloc(dnum * QP * N * sizeof(uint64_t));
    key->B.alloc(dnum * n * QP * N * sizeof(uint64_t));
    const uint32_t *table = c.galois_table(elt, s);
    Scratch sig(n * QP * N * sizeof(uint64_t), s), e(QP * N * sizeof(uint64_t), s);
    for (size_t i = 0; i < n; i++)
        launch_permute(c, K.aux[i].s.as<uint64_t>(), sig.as<uint64_t>() + i * QP * N, table, QP, s);
    const uint64_t kseed = host_splitmix64(seed, elt);
    for (size_t g = 0; g < dnum; g++) {
        uint64_t *Ag = key->A.as<uint64_t>() + g * QP * N;
        launch_sample_uniform(c, host_splitmix64(kseed, g * (n + 1)), Ag, c.nQ, true, s);
        for (size_t i = 0; i < n; i++) {
            launch_sample_gaussian_rns(c, host_splitmix64(kseed, g * (n + 1) + 1 + i), e.as<uint64_t>(), c.nQ, true,
                                       s);
            ntt_forward(c, e.as<uint64_t>(), 1, QP, 0, s);
            launch_ks_key_digit(c, K.aux[i].s.as<uint64_t>(), sig.as<uint64_t>() + i * QP * N,
                                key->B.as<uint64_t>() + (g * n + i) * QP * N, Ag, e.as<uint64_t>(), g, s);
        }
    }
    SHM_HIP(hipStreamSynchronize(s));
    K.galois[elt] = std::move(key);
}

// ---- format switching ---------------------------------------------------------------------

void fmt_switch_forward(SharedAKeys &K, const std::vector<const Ciphertext *> &cts, SharedABatch &out,
                        hipStream_t s, bool rounded) {
    Context &c = *K.ctx;
    const size_t n = K.n, N = c.N, nP = c.nP;
    if (cts.size() != n) throw InvalidArgument("fmt_switch_forward: number of ciphertexts does not match the format key");
    const size_t Ql = c.ql_of_chain(K.fmt_chain), QlP = Ql + nP;
    for (const Ciphertext *ct : cts) {
        if (ct->chain != K.fmt_chain || ct->size != 2)
            throw InvalidArgument("fmt_switch_forward: ciphertext level/size mismatch with the format key");
        if (!ct->ntt) throw InvalidArgument("fmt_switch_forward: ciphertexts must be in NTT form");
        if (std::fabs(ct->scale - cts[0]->scale) > 1e-6 * cts[0]->scale)
            throw InvalidArgument("fmt_switch_forward: all ciphertexts must share the same scale");
    }
    // 1. exact extension of every a_j to QlP (own limbs stay in NTT form)
    Scratch ext(n * QlP * N * sizeof(uint64_t), s), coef(n * Ql * N * sizeof(uint64_t), s);
    for (size_t j = 0; j < n; j++) {
        launch_copy_words(ext.as<uint64_t>() + j * QlP * N, cts[j]->poly(1), Ql * N, s);
        launch_copy_words(coef.as<uint64_t>() + j * Ql * N, cts[j]->poly(1), Ql * N, s);
    }
    ntt_inverse(c, coef.as<uint64_t>(), n, Ql, 0, s);
    k_extend_exact<<<grid_for(n * N), kBlock, 0, s>>>(
        coef.as<uint64_t>(), ext.as<uint64_t>() + Ql * N, c.mods(), K.ext_inv.as<uint64_t>(),
        K.ext_inv_s.as<uint64_t>(), K.ext_cg.as<uint64_t>(), K.ext_qmodp.as<uint64_t>(), Ql, c.nQ, nP, N, n, Ql * N,
        QlP * N);
    SHM_HIP(hipGetLastError());
    ntt_forward_qlp(c, ext.as<uint64_t>(), Ql, Ql, QlP, 0, 0, n, QlP * N, s);
    // 2. inner products: outputs 0..n-1 = b-parts (B[.][i]), output n = a-part (a_fmt)
    KeyRefs kr{};
    for (size_t i = 0; i < n; i++) {
        kr.base[i] = K.B_fmt.as<uint64_t>() + i * QlP * N;
        kr.stride[i] = n * QlP * N;
    }
    kr.base[n] = K.a_fmt.as<uint64_t>();
    kr.stride[n] = QlP * N;
    Scratch cx((n + 1) * QlP * N * sizeof(uint64_t), s);
    ks_inner_multi(c, ext.as<uint64_t>(), Ql, n, kr, n + 1, true, cx.as<uint64_t>(), s);
    // 3. ModDown into the batch: b'_i = b_i + ., a' = .
    out.n = n;
    out.ct.ctx = &c;
    out.ct.resize(n + 1, K.fmt_chain, s);
    out.ct.ntt = true;
    out.ct.scale = cts[0]->scale;
    for (size_t i = 0; i < n; i++)
        launch_copy_words(out.ct.poly(i), cts[i]->poly(0), Ql * N, s);
    const uint64_t mask = (n >= 64) ? ~0ull : ((1ull << n) - 1);
    ks_moddown(c, cx.as<uint64_t>(), n + 1, Ql, out.ct.data(), Ql * N, mask, s, rounded);
}

void fmt_switch_backward(SharedAKeys &K, const SharedABatch &in, const std::vector<Ciphertext *> &out,
                         hipStream_t s, bool rounded) {
    Context &c = *K.ctx;
    const size_t n = in.n, N = c.N;
    if (n != K.n) throw InvalidArgument("fmt_switch_backward: batch size does not match the format key");
    if (out.size() != n) throw InvalidArgument("fmt_switch_backward: output count mismatch");
    if (!in.ct.ntt) throw InvalidArgument("fmt_switch_backward: only NTT form is supported");
    const size_t Ql = in.ct.Ql(), QlP = Ql + c.nP, beta = c.level(Ql).beta;
    Scratch d(beta * QlP * N * sizeof(uint64_t), s);
    ks_modup(c, in.ct.poly(n), Ql, d.as<uint64_t>(), s, rounded);
    Scratch cx(2 * n * QlP * N * sizeof(uint64_t), s);
    // at most kMaxKeys outputs per inner-product launch: members in groups of 16
    for (size_t i0 = 0; i0 < n; i0 += 16) {
        const size_t m = std::min<size_t>(16, n - i0);
        KeyRefs kr{};
        for (size_t t = 0; t < m; t++) {
            const uint64_t *k = K.backward[i0 + t]->k.as<uint64_t>();
            kr.base[2 * t] = k;
            kr.stride[2 * t] = 2 * c.nQP() * N;
            kr.base[2 * t + 1] = k + c.nQP() * N;
            kr.stride[2 * t + 1] = 2 * c.nQP() * N;
        }
        ks_inner_multi(c, d.as<uint64_t>(), Ql, beta, kr, 2 * m, false, cx.as<uint64_t>() + 2 * i0 * QlP * N, s);
    }
    for (size_t i = 0; i < n; i++) {
        Ciphertext &o = *out[i];
        o.ctx = &c;
        o.resize(2, in.ct.chain, s);
        o.ntt = true;
        o.scale = in.ct.scale;
        launch_copy_words(o.poly(0), in.ct.poly(i), Ql * N, s);
        ks_moddown(c, cx.as<uint64_t>() + 2 * i * QlP * N, 2, Ql, o.data(), Ql * N, 1, s, rounded);
    }
}

// ---- shared-a rotation ----------------------------------------------------------------------

void shared_apply_galois(SharedAKeys &K, SharedABatch &x, uint32_t elt, hipStream_t s, bool rounded) {
    Context &c = *K.ctx;
    auto it = K.galois.find(elt);
    if (it == K.galois.end()) throw InvalidArgument("shared-a Galois key not present");
    const SharedGaloisKey &key = *it->second;
    const size_t n = x.n, N = c.N, Ql = x.ct.Ql(), QlP = Ql + c.nP, QP = c.nQP(), beta = c.level(Ql).beta;
    if (x.ct.size != n + 1 || n != K.n) throw InvalidArgument("shared_a_rotate: batch size does not match the Galois key");
    const uint32_t *table = c.galois_table(elt, s);
    if (rounded) {
        // the exact centred ModUp commutes with the Galois permutation: raise the
        // unpermuted a, read its digits through the permutation, and add the members'
        // b parts through it in ModDown's last pass (the hoisted form of one step)
        Scratch d(beta * QlP * N * sizeof(uint64_t), s), cx((n + 1) * QlP * N * sizeof(uint64_t), s);
        ks_modup(c, x.ct.poly(n), Ql, d.as<uint64_t>(), s, true);
        KeyRefs kr{};
        for (size_t i = 0; i < n; i++) {
            kr.base[i] = key.B.as<uint64_t>() + i * QP * N;
            kr.stride[i] = n * QP * N;
        }
        kr.base[n] = key.A.as<uint64_t>();
        kr.stride[n] = QP * N;
        ks_inner_multi(c, d.as<uint64_t>(), Ql, beta, kr, n + 1, false, cx.as<uint64_t>(), s, table);
        Ciphertext o;
        o.ctx = &c;
        o.resize(n + 1, x.ct.chain, s, false);  // the fused ModDown writes every word
        o.scale = x.ct.scale;
        o.ntt = true;
        PolyPack outp{}, srcp{};
        for (size_t p = 0; p <= n; p++) outp.p[p] = o.poly(p);
        for (size_t p = 0; p < n; p++) srcp.p[p] = x.ct.poly(p);
        ks_moddown_pack(c, cx.as<uint64_t>(), n + 1, Ql, outp, (1ull << n) - 1, s, true, &srcp, table);
        x.ct = std::move(o);
        return;
    }
    Ciphertext rot;
    rot.ctx = &c;
    rot.resize(n + 1, x.ct.chain, s);
    rot.ntt = true;
    rot.scale = x.ct.scale;
    launch_permute(c, x.ct.data(), rot.data(), table, (n + 1) * Ql, s);
    Scratch d(beta * QlP * N * sizeof(uint64_t), s), cx((n + 1) * QlP * N * sizeof(uint64_t), s);
    ks_modup(c, rot.poly(n), Ql, d.as<uint64_t>(), s, rounded);
    KeyRefs kr{};
    for (size_t i = 0; i < n; i++) {
        kr.base[i] = key.B.as<uint64_t>() + i * QP * N;
        kr.stride[i] = n * QP * N;
    }
    kr.base[n] = key.A.as<uint64_t>();
    kr.stride[n] = QP * N;
    ks_inner_multi(c, d.as<uint64_t>(), Ql, beta, kr, n + 1, false, cx.as<uint64_t>(), s);
    ks_moddown(c, cx.as<uint64_t>(), n + 1, Ql, rot.data(), Ql * N, (1ull << n) - 1, s, rounded);
    x.ct = std::move(rot);
}

void shared_rotate_hoisted(SharedAKeys &K, const Ciphertext &x, const std::vector<int> &steps,
                           const std::vector<Ciphertext *> &outs, hipStream_t s, bool rounded) {
    Context &c = *K.ctx;
    const size_t n = K.n, N = c.N, Ql = x.Ql(), QlP = Ql + c.nP, QP = c.nQP(), beta = c.level(Ql).beta;
    if (x.size != n + 1) throw InvalidArgument("shared_a_rotate: batch size does not match the Galois key");
    Scratch d(beta * QlP * N * sizeof(uint64_t), s), cx((n + 1) * QlP * N * sizeof(uint64_t), s);
    bool have_modup = false;
    for (size_t t = 0; t < steps.size(); t++) {
        Ciphertext &o = *outs[t];
        const uint32_t elt = steps[t] == 0 ? (uint32_t)(2 * N - 1) : c.elt_from_step(steps[t]);
        auto it = K.galois.find(elt);
        if (it == K.galois.end()) {
            SharedABatch tmp;
            tmp.n = n;
            ct_clone(tmp.ct, x, s);
            shared_rotate(K, tmp, steps[t], s, rounded);
            o = std::move(tmp.ct);
            continue;
        }
        if (!have_modup) {
            ks_modup(c, x.poly(n), Ql, d.as<uint64_t>(), s, rounded);
            have_modup = true;
        }
        const uint32_t *table = c.galois_table(elt, s);
        const SharedGaloisKey &key = *it->second;
        KeyRefs kr{};
        for (size_t i = 0; i < n; i++) {
            kr.base[i] = key.B.as<uint64_t>() + i * QP * N;
            kr.stride[i] = n * QP * N;
        }
        kr.base[n] = key.A.as<uint64_t>();
        kr.stride[n] = QP * N;
        ks_inner_multi(c, d.as<uint64_t>(), Ql, beta, kr, n + 1, false, cx.as<uint64_t>(), s, table);
        o.ctx = &c;
        o.resize(n + 1, x.chain, s, false);  // the fused ModDown writes every word
        o.scale = x.scale;
        o.ntt = true;
        // the members' b parts are added through the Galois permutation by the ModDown's
        // last pass (no permuted copy in between)
        PolyPack outp{}, srcp{};
        for (size_t p = 0; p <= n; p++) outp.p[p] = o.poly(p);
        for (size_t p = 0; p < n; p++) srcp.p[p] = x.poly(p);
        ks_moddown_pack(c, cx.as<uint64_t>(), n + 1, Ql, outp, (1ull << n) - 1, s, rounded, &srcp, table);
    }
}

// shared_a_rotate by a slot step (0 = conjugation); a missing step is split in
// its NAF terms like rotate_internal.
void shared_rotate(SharedAKeys &K, SharedABatch &x, int step, hipStream_t s, bool rounded) {
    Context &c = *K.ctx;
    if (step == 0) {
        shared_apply_galois(K, x, (uint32_t)(2 * c.N - 1), s, rounded);
        return;
    }
    const uint32_t elt = c.elt_from_step(step);
    if (K.galois.count(elt)) {
        shared_apply_galois(K, x, elt, s, rounded);
        return;
    }
    std::vector<int> naf;
    const bool sign = step < 0;
    int v = std::abs(step);
    for (int i = 0; v; i++) {
        const int zi = (v & 1) ? 2 - (v & 3) : 0;
        v = (v - zi) >> 1;
        if (zi) naf.push_back((sign ? -zi : zi) * (1 << i));
    }
    if (naf.size() == 1) throw InvalidArgument("shared-a Galois key not present");
    for (int t : naf)
        if ((size_t)std::abs(t) != (c.N >> 1)) shared_rotate(K, x, t, s, rounded);
}

// ---- batched bootstrapping -------------------------------------------------------------------

namespace {
Rotator shared_rotation(SharedAKeys &K, hipStream_t s) {
    Rotator r;
    r.rot = [&K, s](Ciphertext &x, int step) {
        SharedABatch tmp;
        tmp.n = K.n;
        tmp.ct = std::move(x);
        shared_rotate(K, tmp, step, s, true);
        x = std::move(tmp.ct);
    };
    r.hoist = [&K, s](const Ciphertext &x, const std::vector<int> &steps, const std::vector<Ciphertext *> &outs) {
        shared_rotate_hoisted(K, x, steps, outs, s, true);
    };
    return r;
}
}  // namespace

// shared_a_sfl_full_3 (SharedACoeffToSlot.cu:311-358)
void shared_slot_to_coeff(Bootstrapper &b, SharedAKeys &K, SharedABatch &x, hipStream_t s) {
    if (x.n != K.n) throw InvalidArgument("batch size does not match the shared-a keys");
    if (!b.prepared) b.prepare(s);
    slot_to_coeff(b, x.ct, shared_rotation(K, s), s);
}

// shared_a_coefftoslot_full_3 (SharedACoeffToSlot.cu:269-301)
void shared_coeff_to_slot(Bootstrapper &b, SharedAKeys &K, SharedABatch &x, SharedABatch &re, SharedABatch &im,
                          hipStream_t s) {
    if (x.n != K.n) throw InvalidArgument("batch size does not match the shared-a keys");
    if (!b.prepared) b.prepare(s);
    re.n = im.n = K.n;
    coeff_to_slot(b, x.ct, re.ct, im.ct, shared_rotation(K, s), s);
}

// BatchBootstrap.cuh:719-820 bootstrap_shared_a_s2c_first: FmtSwitch -> Sh-S2C ->
// shared-a modraise -> Sh-C2S -> backward FmtSwitch -> per-member EvalMod +
// recombination (amplitude correction folded into EvalMod).
void bootstrap_shared_a_s2c_first(Bootstrapper &b, SharedAKeys &K, const std::vector<const Ciphertext *> &in,
                                  const std::vector<Ciphertext *> &out, const KSwitchKey &rk, double out_scale,
                                  hipStream_t s) {
    Context &c = *b.ctx;
    if (in.size() != K.n || out.size() != K.n)
        throw InvalidArgument("BatchBootstrapper: shared-a bootstrap needs exactly batch_n ciphertexts");
    if (in[0]->Ql() != b.prm.s2c_input_primes)
        throw InvalidArgument("bootstrap_shared_a_s2c_first: inputs must have exactly 4 primes (lS2C+1)");
    if (K.fmt_chain != in[0]->chain) throw InvalidArgument("format key level does not match the inputs");
    if (!b.prepared) b.prepare(s);
    const double d_in = in[0]->scale;
    if (out_scale <= 0) out_scale = d_in;
    SharedABatch batch;
    fmt_switch_forward(K, in, batch, s, true);
    shared_slot_to_coeff(b, K, batch, s);
    ct_modraise(c, batch.ct, s);
    const double q0 = (double)c.q[0];
    batch.ct.scale = q0;
    SharedABatch re, im;
    shared_coeff_to_slot(b, K, batch, re, im, s);
    const size_t n = K.n;
    std::vector<Ciphertext> res(n), ims(n);
    std::vector<Ciphertext *> pr(n), pi(n);
    for (size_t i = 0; i < n; i++) {
        pr[i] = &res[i];
        pi[i] = &ims[i];
    }
    fmt_switch_backward(K, re, pr, s, true);
    fmt_switch_backward(K, im, pi, s, true);
    const double F = b.prm.layer3_factor;
    const double E = F * d_in * b.prm.input_bound / q0;
    const double pi_ = 3.14159265358979323846;
    const double cst = (1.0 + 0.5 * pi_ * pi_ * E * E) / (2.0 * pi_) * q0 / (F * d_in);
    std::vector<Ciphertext *> parts;
    for (size_t i = 0; i < n; i++) {
        parts.push_back(&res[i]);
        parts.push_back(&ims[i]);
    }
    eval_mod_batch(b, parts, rk, cst, out_scale, s);  // every member's two parts in one batch
    for (size_t i = 0; i < n; i++) {
        ct_mul_monomial_half(c, ims[i], false, s);
        ct_add(c, res[i], ims[i], s);
        res[i].scale = out_scale;
        ct_clone(*out[i], res[i], s);  // the caller's buffer is kept (graph replays write it)
    }
}

}  // namespace boot
}  // namespace shm
