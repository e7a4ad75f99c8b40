// Slot-to-coefficient-first bootstrapping on the device.
//
// Pipeline (BatchBootstrap.cuh:674-709 bootstrap_ordinary_s2c_first):
//   S2C (3 diagonal layers at the 4-prime input level, Bootstrapper.cu:2737-2777)
//   -> modraise to all data primes, scale := q0 (Bootstrapper.cu:3271-3356)
//   -> C2S (3 layers at the top, then w + conj(w) and -i (w - conj(w)),
//      Bootstrapper.cu:3048-3067)
//   -> EvalMod on both halves (ModularReducer.cu:64-84: cosine polynomial,
//      double angles) -> recombine re + i im -> amplitude correction
//      (BatchBootstrap.cuh:444-460).
// Scale bookkeeping is exact: diagonals are encoded at the scale of the prime
// the next rescale divides by, and every constant that meets a ciphertext of
// another scale is encoded at the scale that lands the product on the target.
#include <cmath>

#include "arith.h"
#include "boot.h"

namespace shm {
namespace boot {

namespace {

constexpr int kBlock = 256;
constexpr int kMaxTerms = 16;

// [rows][N] launches: blockIdx.y is the row (one limb of one polynomial), so a
// row's limb is uniform per block; 4 words per thread when the launch is large.
inline dim3 rows_grid(size_t rows, size_t N) {
    size_t bx = (N + kBlock - 1) / kBlock;
    if (rows * bx >= 8192) bx = (bx + 3) / 4;
    const size_t by = rows == 0 ? 1 : (rows < 65535 ? rows : 65535);
    return dim3((unsigned)bx, (unsigned)by);
}

// [sum over k < nterms of x(k) * y(k)]_q for residues x, y < q and nterms <= 16, the
// operands of up to 8 terms loaded before their products so their loads are in flight
// together (nterms and p52 are uniform over the block). With p52 (every prime in
// (2^45, 2^52)) the products are 26-bit split multiply-adds into an Acc52 (sum below
// 16 * 2^104 < 2^109) reduced by reduce52, else 128-bit products and Barrett; the same
// canonical residue either way.
template <typename LX, typename LY>
__device__ __forceinline__ uint64_t sum_terms(int nterms, int p52, const DevMod &m, LX x, LY y) {
    Acc128 a{0, 0};
    Acc52 b{0, 0, 0};
    for (int k = 0; k < nterms; k += 8) {
        uint64_t xv[8], yv[8];
#pragma unroll
        for (int j = 0; j < 8; j++)
            if (k + j < nterms) {
                xv[j] = x(k + j);
                yv[j] = y(k + j);
            }
        if (p52) {
#pragma unroll
            for (int j = 0; j < 8; j++)
                if (k + j < nterms) mac52(b, split26(xv[j]), yv[j]);
        } else {
#pragma unroll
            for (int j = 0; j < 8; j++)
                if (k + j < nterms) mac128(a, xv[j], yv[j]);
        }
    }
    return p52 ? reduce52(b, m) : reduce128(a, m);
}

struct PtPack {
    const uint64_t *x[kMaxTerms];
    const uint64_t *pt[kMaxTerms];
};

// acc[p][l][i] = sum_t x_t[p][l][i] * pt_t[l][i] mod q_l over npolys polynomials;
// products accumulate in 128 bits (at most 16 terms below 2^122).
__global__ void k_lt_mac(uint64_t *acc, PtPack pk, int nterms, const DevMod *mods, size_t Ql, size_t N,
                         size_t npolys, int p52) {
    for (size_t row = blockIdx.y; row < npolys * Ql; row += gridDim.y) {
        const size_t limb = row % Ql;
        const DevMod m = mods[limb];
        for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < N; i += (size_t)gridDim.x * blockDim.x) {
            const uint64_t v = sum_terms(nterms, p52, m, [&](int k) { return pk.x[k][row * N + i]; },
                                         [&](int k) { return pk.pt[k][limb * N + i]; });
            __builtin_nontemporal_store(v, &acc[row * N + i]);
        }
    }
}

struct SrcPack {
    const uint64_t *x[kMaxTerms];
    uint32_t ql[kMaxTerms];  // limb count (stride) of each source ciphertext
    double v[kMaxTerms];     // integral constants round(value * sigma)
};

// acc[p][l][i] = sum_k x_k[p][l][i] * [v_k]_{q_l}: a linear combination of
// ciphertexts at different levels, each read over the first Ql limbs. One row
// (p, l) per blockIdx.y; the constant residues of the row are formed once, from
// the integral doubles, into LDS.
__global__ void k_const_lincomb(uint64_t *acc, SrcPack sp, int nterms, const DevMod *mods, size_t Ql, size_t N,
                                int p52) {
    __shared__ uint64_t cres[kMaxTerms];
    for (size_t row = blockIdx.y; row < 2 * Ql; row += gridDim.y) {
        const size_t p = row / Ql, l = row % Ql;
        const DevMod m = mods[l];
        __syncthreads();  // the previous row's readers are done with cres
        if (threadIdx.x < (unsigned)nterms) {
            const uint64_t c = residue_of_rounded(sp.v[threadIdx.x], m);
            cres[threadIdx.x] = c >= m.q ? c - m.q : c;
        }
        __syncthreads();
        for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < N; i += (size_t)gridDim.x * blockDim.x) {
            const uint64_t v = sum_terms(nterms, p52, m, [&](int k) { return sp.x[k][(p * sp.ql[k] + l) * N + i]; },
                                         [&](int k) { return cres[k]; });
            __builtin_nontemporal_store(v, &acc[row * N + i]);
        }
    }
}

// k_const_lincomb over up to 8 members sharing the constants and source levels:
// row r is (member r / (2 Ql), polynomial (r / Ql) % 2, limb r % Ql).
constexpr int kLeafMembers = 8, kLeafTerms = 8;
struct LeafPack {
    const uint64_t *x[kLeafMembers][kLeafTerms];
    uint64_t *acc[kLeafMembers];
    uint32_t ql[kLeafTerms];
    double v[kLeafTerms];
};

__global__ void k_const_lincomb_batch(LeafPack lp, int nterms, int members, const DevMod *mods, size_t Ql, size_t N,
                                      int p52) {
    __shared__ uint64_t cres[kLeafTerms];
    for (size_t row = blockIdx.y; row < (size_t)members * 2 * Ql; row += gridDim.y) {
        const size_t mb = row / (2 * Ql), p = (row / Ql) % 2, l = row % Ql;
        const DevMod m = mods[l];
        __syncthreads();  // the previous row's readers are done with cres
        if (threadIdx.x < (unsigned)nterms) {
            const uint64_t c = residue_of_rounded(lp.v[threadIdx.x], m);
            cres[threadIdx.x] = c >= m.q ? c - m.q : c;
        }
        __syncthreads();
        uint64_t *out = lp.acc[mb] + (p * Ql + l) * N;
        for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < N; i += (size_t)gridDim.x * blockDim.x) {
            const uint64_t v = sum_terms(nterms, p52, m, [&](int k) { return lp.x[mb][k][(p * lp.ql[k] + l) * N + i]; },
                                         [&](int k) { return cres[k]; });
            __builtin_nontemporal_store(v, &out[i]);
        }
    }
}

void add_checked(Context &c, Ciphertext &a, const Ciphertext &b, hipStream_t s) {
    if (std::fabs(a.scale / b.scale - 1.0) > 1e-9) throw LogicError("scale mismatch");
    ct_add(c, a, b, s);
}

// a and b to a common (level, scale): the operand with more primes is aligned
// onto the other one.
void add_aligned(Context &c, Ciphertext &a, Ciphertext &b, hipStream_t s) {
    if (a.chain < b.chain)
        ct_align(c, a, b.chain, b.scale, s, true);
    else if (b.chain < a.chain)
        ct_align(c, b, a.chain, a.scale, s, true);
    add_checked(c, a, b, s);
}

}  // namespace

void const_lincomb(Context &c, Ciphertext &acc, const std::vector<const Ciphertext *> &src,
                   const std::vector<double> &values, const std::vector<double> &sigmas, size_t Ql, hipStream_t s) {
    if (src.empty() || src.size() > (size_t)kMaxTerms || values.size() != src.size() || sigmas.size() != src.size())
        throw InvalidArgument("const_lincomb: 1..16 terms");
    SrcPack sp{};
    for (size_t k = 0; k < src.size(); k++) {
        if (src[k]->Ql() < Ql || src[k]->size != 2) throw LogicError("const_lincomb: source below the target level");
        sp.x[k] = src[k]->data();
        sp.ql[k] = (uint32_t)src[k]->Ql();
        sp.v[k] = std::round(values[k] * sigmas[k]);
        if (!(std::fabs(sp.v[k]) < 3.4e38)) throw InvalidArgument("constant out of range");
    }
    acc.ctx = &c;
    acc.resize(2, c.nQ - Ql + 1, s, false);  // k_const_lincomb writes every word
    acc.ntt = true;
    k_const_lincomb<<<rows_grid(2 * Ql, c.N), kBlock, 0, s>>>(acc.data(), sp, (int)src.size(), c.mods(), Ql, c.N,
                                                              c.primes52 ? 1 : 0);
    SHM_HIP(hipGetLastError());
}

// bsgs_linear_transform (Bootstrapper.cu:2168-2251): babies rot(x, j*base),
// inner sums with pre-rotated diagonals, giant rotation of each inner sum. `ct`
// is an ordinary ciphertext or a shared-a batch (every polynomial is multiplied
// by the same diagonal); `rot` rotates it in place.
void lt_apply(Context &c, const LtLayer &L, Ciphertext &ct, const Rotator &r, hipStream_t s) {
    if (ct.size < 2) throw InvalidArgument("linear transform takes ciphertexts of size >= 2");
    if (ct.chain != L.chain) throw InvalidArgument("linear transform applied at the wrong level");
    if (L.pts.empty()) throw LogicError("bootstrapper is not prepared");
    const size_t Ql = ct.Ql(), n = c.N / 2;
    std::vector<Ciphertext> baby(L.g);
    std::vector<int> steps;
    std::vector<Ciphertext *> outs;
    for (int j = 1; j < L.g; j++) {
        bool used = false;
        for (int i = 0; i < L.n_giants; i++) used |= !L.pre[i][j].empty();
        if (!used) continue;
        steps.push_back(L.baby_step(j));
        outs.push_back(&baby[j]);
    }
    if (r.hoist) {
        r.hoist(ct, steps, outs);
    } else {
        for (size_t t = 0; t < steps.size(); t++) {
            ct_clone(*outs[t], ct, s);
            r.rot(*outs[t], steps[t]);
        }
    }
    Ciphertext out;
    bool have = false;
    for (int i = 0; i < L.n_giants; i++) {
        PtPack pk{};
        int nt = 0;
        for (int j = 0; j < L.g; j++) {
            if (L.pre[i][j].empty()) continue;
            pk.x[nt] = j == 0 ? ct.data() : baby[j].data();
            pk.pt[nt] = L.pts[i][j].as<uint64_t>();
            nt++;
        }
        if (!nt) continue;
        Ciphertext inner;
        inner.ctx = &c;
        inner.resize(ct.size, ct.chain, s, false);  // k_lt_mac writes every word
        inner.ntt = true;
        inner.scale = ct.scale;
        k_lt_mac<<<rows_grid(ct.size * Ql, c.N), kBlock, 0, s>>>(inner.data(), pk, nt, c.mods(), Ql, c.N, ct.size,
                                                                 c.primes52 ? 1 : 0);
        SHM_HIP(hipGetLastError());
        const long G = ((L.giant_offset(i) % (long)n) + (long)n) % (long)n;
        if (G) r.rot(inner, (int)G);
        if (!have) {
            out = std::move(inner);
            have = true;
        } else {
            ct_add(c, out, inner, s);
        }
    }
    if (!have) throw LogicError("empty linear transform");
    const double sc = ct.scale;
    out.scale = sc * (double)c.q[Ql - 1];
    rescale_inplace(c, out, s, true);
    out.scale = sc;  // diagonals carry exactly the scale q_last
    ct = std::move(out);
}

void slot_to_coeff(Bootstrapper &b, Ciphertext &ct, const Rotator &r, hipStream_t s) {
    for (int l = 0; l < 3; l++) lt_apply(*b.ctx, b.s2c[l], ct, r, s);
}

namespace {
Rotator ordinary_rotator(Context &c, const GaloisKeys &gk, hipStream_t s) {
    Rotator r;
    r.rot = [&c, &gk, s](Ciphertext &x, int step) {
        if (step == 0)
            apply_galois(c, x, (uint32_t)(2 * c.N - 1), gk, s, true);
        else
            rotate_internal(c, x, step, gk, s, true);
    };
    r.hoist = [&c, &gk, s](const Ciphertext &x, const std::vector<int> &steps, const std::vector<Ciphertext *> &outs) {
        rotate_hoisted(c, x, steps, gk, outs, s, true);
    };
    return r;
}

// Rotations of a group container (b_0, a_0, b_1, a_1, ...) of ordinary
// ciphertexts: the linear transforms then run once for the whole group (every
// polynomial meets the same diagonal), with batched key switches.
Rotator group_rotator(Context &c, const GaloisKeys &gk, hipStream_t s) {
    Rotator r;
    r.rot = [&c, &gk, s](Ciphertext &x, int step) {
        if (step == 0)
            apply_galois_group(c, x, (uint32_t)(2 * c.N - 1), gk, s, true);
        else
            rotate_internal_group(c, x, step, gk, s, true);
    };
    r.hoist = [&c, &gk, s](const Ciphertext &x, const std::vector<int> &steps, const std::vector<Ciphertext *> &outs) {
        rotate_hoisted_group(c, x, steps, gk, outs, s, true);
    };
    return r;
}
}  // namespace

void slot_to_coeff(Bootstrapper &b, Ciphertext &ct, const GaloisKeys &gk, hipStream_t s) {
    slot_to_coeff(b, ct, ordinary_rotator(*b.ctx, gk, s), s);
}

// coefftoslot_full_3 (Bootstrapper.cu:3048-3067): re = w + conj(w) and
// im = -i (w - conj(w)) with one conjugation; -i is the exact monomial X^(3N/2).
void coeff_to_slot(Bootstrapper &b, Ciphertext &ct, Ciphertext &re, Ciphertext &im, const Rotator &r,
                   hipStream_t s) {
    Context &c = *b.ctx;
    for (int l = 0; l < 3; l++) lt_apply(c, b.c2s[l], ct, r, s);
    Ciphertext conj;
    ct_clone(conj, ct, s);
    r.rot(conj, 0);  // step 0 = complex conjugation
    ct_clone(re, ct, s);
    ct_add(c, re, conj, s);  // w + conj(w)
    ct_clone(im, ct, s);
    ct_sub(c, im, conj, s);
    ct_mul_monomial_half(c, im, true, s);  // -i (w - conj(w)) = 2 Im(w)
}

void coeff_to_slot(Bootstrapper &b, Ciphertext &ct, Ciphertext &re, Ciphertext &im, const GaloisKeys &gk,
                   hipStream_t s) {
    coeff_to_slot(b, ct, re, im, ordinary_rotator(*b.ctx, gk, s), s);
}

namespace {

struct ChebEval {
    Context &c;
    const KSwitchKey &rk;
    hipStream_t s;
    std::vector<std::unique_ptr<Ciphertext>> T;

    ChebEval(Context &c_, const KSwitchKey &rk_, hipStream_t s_, Ciphertext &&t1) : c(c_), rk(rk_), s(s_), T(65) {
        T[1] = std::make_unique<Ciphertext>(std::move(t1));
    }

    // leaves are combinations of T_1..T_8; a longer polynomial of degree d splits
    // as q + T_m r with m = split_a(d), the largest power of two below d
    static constexpr size_t kLeaf = 8;

    static int split_a(int k) {
        int a = 1;
        while (2 * a < k) a *= 2;
        return a;
    }
    // multiplicative depth of T_k above T_1 under the rule used by get()
    static int depth(int k) {
        if (k <= 1) return 0;
        const int a = split_a(k);
        return std::max(depth(a), depth(k - a)) + 1;
    }
    size_t tchain(int k) const { return T[1]->chain + depth(k); }

    // T_k = 2 T_a T_b - T_{a-b}, a the largest power of two below k (T_{2a} = 2 T_a^2 - 1).
    // The doubling and T_{a-b} (at the product's scale through an integer constant)
    // enter the tensor before the relinearisation, so T_{a-b} needs no rescale of its own.
    Ciphertext &get(int k) {
        if (T[k]) return *T[k];
        const int a = split_a(k), b = k - a;
        Ciphertext &ta = get(a);
        Ciphertext &tb = get(b);
        const std::vector<const Ciphertext *> term{a == b ? nullptr : &get(a - b)};
        auto out = std::make_unique<Ciphertext>();
        mul_relin_rescale_batch(c, {out.get()}, {&ta}, {&tb}, rk, s, true, &term);
        if (a == b) ct_add_const(c, *out, -1.0, s);
        if (out->chain != tchain(k)) throw LogicError("Chebyshev basis level bookkeeping");
        T[k] = std::move(out);
        return *T[k];
    }

    size_t baby_chain(const std::vector<double> &p) const {
        size_t ch = 0;
        for (size_t k = 1; k < p.size(); k++)
            if (p[k] != 0.0) ch = std::max(ch, tchain((int)k));
        return ch + 1;
    }

    // chain index eval() lands on
    size_t eval_chain(const std::vector<double> &p) const {
        const size_t deg = p.size() - 1;
        if (deg <= kLeaf) return baby_chain(p);
        const size_t m = (size_t)split_a((int)deg);
        std::vector<double> q, r;
        cheb_split(p, m, q, r);
        return std::max(tchain((int)m), eval_chain(r)) + 1;
    }

    // sum_k p[k] T_k landing exactly at (chain, scale): products at chain-1 with
    // each constant encoded at scale * q_last / scale(T_k), one rescale.
    void baby(const std::vector<double> &p, size_t chain, double scale, Ciphertext &out) {
        std::vector<const Ciphertext *> src;
        std::vector<double> vals, sig;
        const size_t Qc = c.ql_of_chain(chain - 1);
        const double ql = (double)c.q[Qc - 1];
        for (size_t k = 1; k < p.size(); k++) {
            if (p[k] == 0.0) continue;
            Ciphertext &t = get((int)k);
            if (t.chain > chain - 1) throw LogicError("Chebyshev basis element below the combination level");
            src.push_back(&t);
            vals.push_back(p[k]);
            sig.push_back(scale * ql / t.scale);
        }
        if (src.empty()) throw LogicError("constant-only Chebyshev block");
        const_lincomb(c, out, src, vals, sig, Qc, s);
        out.scale = scale * ql;
        rescale_inplace(c, out, s, true);
        out.scale = scale;
        if (p[0] != 0.0) ct_add_const(c, out, p[0], s);
    }

    // Homomorphic_poly_evaluation (common/Polynomial.cu:253-476) in Chebyshev
    // form: p = q + T_m r with m = split_a(deg) in {8, 16, 32}, recursively, down
    // to leaves over the babies T_1..T_8 (degree 63: 7 baby and 2 giant products,
    // 7 combination products, depth 7). The result carries exactly `target` as
    // its scale: r is evaluated at the scale that the product with T_m maps onto
    // `target`.
    void eval(const std::vector<double> &p, double target, Ciphertext &out) {
        const size_t deg = p.size() - 1;
        if (deg <= kLeaf) {
            baby(p, baby_chain(p), target, out);
            return;
        }
        const size_t m = (size_t)split_a((int)deg);
        std::vector<double> q, r;
        cheb_split(p, m, q, r);
        const size_t pc = std::max(tchain((int)m), eval_chain(r));
        const double qp = (double)c.q[c.ql_of_chain(pc) - 1];
        Ciphertext &Tm = get((int)m);
        Ciphertext R;
        eval(r, target * qp / Tm.scale, R);
        ct_mul_relin_rescale(c, out, Tm, R, rk, s, true);
        out.scale = target;
        bool q_terms = false;
        for (size_t k = 1; k < q.size(); k++) q_terms |= q[k] != 0.0;
        if (!q_terms) {
            if (q[0] != 0.0) ct_add_const(c, out, q[0], s);
            return;
        }
        Ciphertext Q;
        if (q.size() <= kLeaf + 1) {
            baby(q, out.chain, target, Q);
        } else {
            eval(q, target, Q);
            if (Q.chain < out.chain) ct_align(c, Q, out.chain, target, s, true);
        }
        add_checked(c, out, Q, s);
    }
};

// ChebEval over m ciphertexts that go through the same polynomial in lockstep:
// every product of the recursion is one mul_relin_rescale_batch for the group
// and every rescale one rescale_batch, so each stage is a handful of launches
// for all members and the relinearisation key is streamed once per group.
// Member i sees exactly the operations ChebEval applies to it alone.
struct ChebEvalBatch {
    typedef std::vector<std::unique_ptr<Ciphertext>> Group;
    Context &c;
    const KSwitchKey &rk;
    hipStream_t s;
    size_t m;
    std::vector<Group> T;

    ChebEvalBatch(Context &c_, const KSwitchKey &rk_, hipStream_t s_, Group &&t1)
        : c(c_), rk(rk_), s(s_), m(t1.size()), T(65) {
        T[1] = std::move(t1);
    }

    static std::vector<Ciphertext *> ptrs(const Group &g) {
        std::vector<Ciphertext *> v;
        for (auto &p : g) v.push_back(p.get());
        return v;
    }
    static std::vector<const Ciphertext *> cptrs(const Group &g) {
        std::vector<const Ciphertext *> v;
        for (auto &p : g) v.push_back(p.get());
        return v;
    }
    Group fresh() const {
        Group g(m);
        for (auto &p : g) p = std::make_unique<Ciphertext>();
        return g;
    }
    size_t tchain(int k) const { return T[1][0]->chain + ChebEval::depth(k); }

    // ct_align of every member onto (chain, scale), one rescale pass for the group
    void align(std::vector<Ciphertext *> xs, size_t chain, double scale) {
        std::vector<Ciphertext *> todo;
        for (Ciphertext *x : xs) {
            if (x->chain > chain) throw LogicError("ct_align: ciphertext is already below the target level");
            if (x->chain == chain) {
                if (std::fabs(x->scale / scale - 1.0) > 1e-9) throw LogicError("ct_align: scales differ at equal levels");
                x->scale = scale;
                continue;
            }
            if (x->chain + 1 < chain) mod_switch_drop(c, *x, chain - 1, s);
            const double qlast = (double)c.q[x->Ql() - 1];
            ct_mul_const(c, *x, 1.0, scale * qlast / x->scale, s);
            todo.push_back(x);
        }
        rescale_batch(c, todo, s, true);
        for (Ciphertext *x : todo) x->scale = scale;
    }

    const Group &get(int k) {
        if (!T[k].empty()) return T[k];
        const int a = ChebEval::split_a(k), b = k - a;
        const Group &ta = get(a);
        const Group &tb = get(b);
        std::vector<const Ciphertext *> term(m, nullptr);
        if (a != b) term = cptrs(get(a - b));
        Group out = fresh();
        mul_relin_rescale_batch(c, ptrs(out), cptrs(ta), cptrs(tb), rk, s, true, &term);
        if (a == b) ct_add_const_batch(c, ptrs(out), -1.0, s);
        if (out[0]->chain != tchain(k)) throw LogicError("Chebyshev basis level bookkeeping");
        T[k] = std::move(out);
        return T[k];
    }

    size_t baby_chain(const std::vector<double> &p) const {
        size_t ch = 0;
        for (size_t k = 1; k < p.size(); k++)
            if (p[k] != 0.0) ch = std::max(ch, tchain((int)k));
        return ch + 1;
    }
    size_t eval_chain(const std::vector<double> &p) const {
        const size_t deg = p.size() - 1;
        if (deg <= ChebEval::kLeaf) return baby_chain(p);
        const size_t mm = (size_t)ChebEval::split_a((int)deg);
        std::vector<double> q, r;
        cheb_split(p, mm, q, r);
        return std::max(tchain((int)mm), eval_chain(r)) + 1;
    }

    void baby(const std::vector<double> &p, size_t chain, double scale, const Group &out) {
        const size_t Qc = c.ql_of_chain(chain - 1);
        const double ql = (double)c.q[Qc - 1];
        for (size_t k = 1; k < p.size(); k++)
            if (p[k] != 0.0) get((int)k);
        // one launch per 8 members: the terms' constants and levels are the group's
        std::vector<int> ks;
        LeafPack lp{};
        for (size_t k = 1; k < p.size(); k++) {
            if (p[k] == 0.0) continue;
            const Ciphertext &t = *T[k][0];
            if (t.chain > chain - 1) throw LogicError("Chebyshev basis element below the combination level");
            if (ks.size() == (size_t)kLeafTerms) throw LogicError("Chebyshev leaf with more than 8 terms");
            lp.ql[ks.size()] = (uint32_t)t.Ql();
            lp.v[ks.size()] = std::round(p[k] * (scale * ql / t.scale));
            if (!(std::fabs(lp.v[ks.size()]) < 3.4e38)) throw InvalidArgument("constant out of range");
            ks.push_back((int)k);
        }
        if (ks.empty()) throw LogicError("constant-only Chebyshev block");
        for (size_t i0 = 0; i0 < m; i0 += kLeafMembers) {
            const size_t mm = std::min<size_t>(kLeafMembers, m - i0);
            for (size_t i = 0; i < mm; i++) {
                Ciphertext &o = *out[i0 + i];
                o.ctx = &c;
                o.resize(2, c.nQ - Qc + 1, s, false);  // every word is written
                o.ntt = true;
                o.scale = scale * ql;
                lp.acc[i] = o.data();
                for (size_t j = 0; j < ks.size(); j++) {
                    const Ciphertext &t = *T[ks[j]][i0 + i];
                    if (t.chain != T[ks[j]][0]->chain || t.size != 2 || t.scale != T[ks[j]][0]->scale)
                        throw LogicError("Chebyshev basis members at different levels or scales");
                    lp.x[i][j] = t.data();
                }
            }
            k_const_lincomb_batch<<<rows_grid(2 * Qc * mm, c.N), kBlock, 0, s>>>(lp, (int)ks.size(), (int)mm, c.mods(),
                                                                                Qc, c.N, c.primes52 ? 1 : 0);
            SHM_HIP(hipGetLastError());
        }
        rescale_batch(c, ptrs(out), s, true);
        for (auto &o : out) o->scale = scale;
        if (p[0] != 0.0) ct_add_const_batch(c, ptrs(out), p[0], s);
    }

    void eval(const std::vector<double> &p, double target, const Group &out) {
        const size_t deg = p.size() - 1;
        if (deg <= ChebEval::kLeaf) {
            baby(p, baby_chain(p), target, out);
            return;
        }
        const size_t mm = (size_t)ChebEval::split_a((int)deg);
        std::vector<double> q, r;
        cheb_split(p, mm, q, r);
        const size_t pc = std::max(tchain((int)mm), eval_chain(r));
        const double qp = (double)c.q[c.ql_of_chain(pc) - 1];
        const Group &Tm = get((int)mm);
        Group R = fresh();
        eval(r, target * qp / Tm[0]->scale, R);
        mul_relin_rescale_batch(c, ptrs(out), cptrs(Tm), cptrs(R), rk, s, true);
        for (auto &o : out) o->scale = target;
        bool q_terms = false;
        for (size_t k = 1; k < q.size(); k++) q_terms |= q[k] != 0.0;
        if (!q_terms) {
            if (q[0] != 0.0) ct_add_const_batch(c, ptrs(out), q[0], s);
            return;
        }
        Group Q = fresh();
        if (q.size() <= ChebEval::kLeaf + 1) {
            baby(q, out[0]->chain, target, Q);
        } else {
            eval(q, target, Q);
            std::vector<Ciphertext *> low;
            for (auto &x : Q)
                if (x->chain < out[0]->chain) low.push_back(x.get());
            align(low, out[0]->chain, target);
        }
        for (size_t i = 0; i < m; i++)
            if (std::fabs(out[i]->scale / Q[i]->scale - 1.0) > 1e-9) throw LogicError("scale mismatch");
        ct_add_batch(c, ptrs(out), cptrs(Q), s);
    }
};

}  // namespace

// modular_reduction of several ciphertexts at once (same level and scale, up to
// 32): the batched form of eval_mod below, bit-identical to it member by member
// (BatchBootstrap.cuh:192 ordinary_evalmod_combine_batch runs EvalMod per chunk).
void eval_mod_batch(Bootstrapper &b, const std::vector<Ciphertext *> &xs, const KSwitchKey &rk, double out_const,
                    double out_scale, hipStream_t s) {
    Context &c = *b.ctx;
    const size_t m = xs.size();
    if (m == 0) return;
    if (m > 32) throw InvalidArgument("eval_mod_batch: at most 32 ciphertexts");
    const int r = b.prm.double_angles;
    ChebEvalBatch::Group t1(m);
    for (size_t i = 0; i < m; i++) {
        if (xs[i]->chain != xs[0]->chain || xs[i]->scale != xs[0]->scale)
            throw InvalidArgument("eval_mod_batch: inputs at different levels or scales");
        t1[i] = std::make_unique<Ciphertext>();
        ct_clone(*t1[i], *xs[i], s);
    }
    const double s1 = xs[0]->scale;
    ct_mul_const_batch(c, ChebEvalBatch::ptrs(t1), 1.0 / (double)b.prm.K,
                       s1 * (double)c.q[t1[0]->Ql() - 1] / t1[0]->scale, s);
    rescale_batch(c, ChebEvalBatch::ptrs(t1), s, true);
    for (auto &t : t1) t->scale = s1;
    ChebEvalBatch ev(c, rk, s, std::move(t1));
    std::vector<double> coeffs = b.cheb;
    if (r == 0)
        for (auto &v : coeffs) v *= out_const;
    const size_t cp = ev.eval_chain(coeffs);
    if (c.ql_of_chain(cp) < (size_t)r + 1) throw LogicError("end of modulus switching chain reached");
    std::vector<double> sig(r + 1);
    sig[r] = out_scale;
    for (int j = r - 1; j >= 0; j--) {
        const double qj = (double)c.q[c.ql_of_chain(cp + j) - 1];
        sig[j] = std::sqrt(2.0 * (j == r - 1 ? out_const : 1.0) * qj * sig[j + 1]);
    }
    ChebEvalBatch::Group out = ev.fresh();
    ev.eval(coeffs, sig[0], out);
    if (out[0]->chain != cp) throw LogicError("EvalMod level bookkeeping");
    for (int j = 0; j < r; j++) {
        ChebEvalBatch::Group sq = ev.fresh();
        mul_relin_rescale_batch(c, ChebEvalBatch::ptrs(sq), ChebEvalBatch::cptrs(out), ChebEvalBatch::cptrs(out), rk,
                                s, true);
        const double cst = j == r - 1 ? out_const : 1.0;
        for (auto &x : sq) x->scale = sig[j + 1];  // = sig[j]^2 / (2 cst q_j): the square read as 2 cst c^2
        ct_add_const_batch(c, ChebEvalBatch::ptrs(sq), -cst, s);
        out = std::move(sq);
    }
    for (size_t i = 0; i < m; i++) *xs[i] = std::move(*out[i]);
}

void chebyshev_eval_batch(Context &c, const std::vector<Ciphertext *> &xs, const std::vector<double> &coeffs,
                          double lo, double hi, const KSwitchKey &rk, hipStream_t s) {
    const size_t m = xs.size();
    if (m == 0 || m > 32) throw InvalidArgument("evaluate_chebyshev: 1..32 ciphertexts");
    if (coeffs.size() < 2 || coeffs.size() > 64) throw InvalidArgument("evaluate_chebyshev: degree 1..63");
    bool any = false;
    for (size_t k = 1; k < coeffs.size(); k++) any |= coeffs[k] != 0.0;
    if (!any) throw InvalidArgument("evaluate_chebyshev: every non-constant coefficient is zero");
    if (!(hi > lo)) throw InvalidArgument("evaluate_chebyshev: empty interval");
    ChebEvalBatch::Group t1(m);
    for (size_t i = 0; i < m; i++) {
        if (xs[i]->size != 2) throw InvalidArgument("evaluate_chebyshev takes size-2 ciphertexts");
        if (xs[i]->chain != xs[0]->chain || xs[i]->scale != xs[0]->scale)
            throw InvalidArgument("evaluate_chebyshev: inputs at different levels or scales");
        t1[i] = std::make_unique<Ciphertext>();
        ct_clone(*t1[i], *xs[i], s);
    }
    // u = (2 x - (lo + hi)) / (hi - lo): one exact constant multiply and rescale
    // back onto the input scale, then the constant term
    const double s1 = xs[0]->scale;
    if (c.ql_of_chain(xs[0]->chain) < 2) throw LogicError("end of modulus switching chain reached");
    ct_mul_const_batch(c, ChebEvalBatch::ptrs(t1), 2.0 / (hi - lo), s1 * (double)c.q[t1[0]->Ql() - 1] / t1[0]->scale,
                       s);
    rescale_batch(c, ChebEvalBatch::ptrs(t1), s, true);
    for (auto &t : t1) t->scale = s1;
    const double shift = -(lo + hi) / (hi - lo);
    if (shift != 0.0) ct_add_const_batch(c, ChebEvalBatch::ptrs(t1), shift, s);
    ChebEvalBatch ev(c, rk, s, std::move(t1));
    const size_t cp = ev.eval_chain(coeffs);
    if (cp > c.nQ) throw LogicError("end of modulus switching chain reached");
    ChebEvalBatch::Group out = ev.fresh();
    ev.eval(coeffs, s1, out);
    if (out[0]->chain != cp) throw LogicError("Chebyshev level bookkeeping");
    for (size_t i = 0; i < m; i++) *xs[i] = std::move(*out[i]);
}

// ModularReducer::modular_reduction (ModularReducer.cu:64-84). x holds slots
// x' = eps + integer with |x'| < K at any scale; the result holds
// out_const * sin(2 pi x') = out_const * cos(2 pi (x' - 1/4)) at scale out_scale.
// Steps: y = x'/K (one constant multiply, so that the noise keeps its size
// relative to x'), the Chebyshev polynomial of cos(2 pi (K y - 1/4) / 2^r), then
// r double angles c -> 2c^2 - 1. Each double angle reads its square at half the
// scale (2c^2 for free) and the last one also absorbs out_const; the polynomial
// targets the scale that this chain of squarings maps onto out_scale.
void eval_mod(Bootstrapper &b, Ciphertext &x, const KSwitchKey &rk, double out_const, double out_scale,
              hipStream_t s) {
    Context &c = *b.ctx;
    const int r = b.prm.double_angles;
    Ciphertext t1;
    ct_clone(t1, x, s);
    const double s1 = x.scale;
    ct_mul_const_rescale(c, t1, 1.0 / (double)b.prm.K, s1 * (double)c.q[t1.Ql() - 1] / t1.scale, s, true);
    t1.scale = s1;
    ChebEval ev(c, rk, s, std::move(t1));
    std::vector<double> coeffs = b.cheb;
    if (r == 0)
        for (auto &v : coeffs) v *= out_const;
    const size_t cp = ev.eval_chain(coeffs);
    if (c.ql_of_chain(cp) < (size_t)r + 1) throw LogicError("end of modulus switching chain reached");
    std::vector<double> sig(r + 1);
    sig[r] = out_scale;
    for (int j = r - 1; j >= 0; j--) {
        const double qj = (double)c.q[c.ql_of_chain(cp + j) - 1];
        sig[j] = std::sqrt(2.0 * (j == r - 1 ? out_const : 1.0) * qj * sig[j + 1]);
    }
    Ciphertext out;
    ev.eval(coeffs, sig[0], out);
    if (out.chain != cp) throw LogicError("EvalMod level bookkeeping");
    for (int j = 0; j < r; j++) {
        Ciphertext sq;
        ct_mul_relin_rescale(c, sq, out, out, rk, s, true);
        const double cst = j == r - 1 ? out_const : 1.0;
        sq.scale = sig[j + 1];  // = sig[j]^2 / (2 cst q_j): the square read as 2 cst c^2
        ct_add_const(c, sq, -cst, s);
        out = std::move(sq);
    }
    x = std::move(out);
}

void bootstrap_s2c_first(Bootstrapper &b, const Ciphertext &in, Ciphertext &out, const GaloisKeys &gk,
                         const KSwitchKey &rk, double out_scale, hipStream_t s) {
    bootstrap_s2c_first_batch(b, {&in}, {&out}, gk, rk, out_scale, s);
}

void bootstrap_s2c_first_batch(Bootstrapper &b, const std::vector<const Ciphertext *> &in,
                               const std::vector<Ciphertext *> &out, const GaloisKeys &gk, const KSwitchKey &rk,
                               double out_scale, hipStream_t s) {
    Context &c = *b.ctx;
    if (in.size() != out.size()) throw InvalidArgument("bootstrap: input and output counts differ");
    for (const Ciphertext *x : in) {
        if (x->size != 2) throw InvalidArgument("bootstrapping takes size-2 ciphertexts");
        if (x->Ql() != b.prm.s2c_input_primes)
            throw InvalidArgument("bootstrap_s2c_first: inputs must have exactly 4 primes (lS2C+1)");
    }
    if (!b.prepared) b.prepare(s);
    const double q0 = (double)c.q[0];
    for (size_t i = 0; i < in.size();) {
        size_t j = i + 1;
        // groups of up to 8: measured at C3, 1 / 2 / 4 / 8 / 16 per group give
        // 57.5 / 48.2 / 43.8 / 42.7 / 55 ms per ciphertext
        while (j < in.size() && j - i < 8 && in[j]->scale == in[i]->scale) j++;
        const size_t g = j - i;
        const double d_in = in[i]->scale, osc = out_scale > 0 ? out_scale : d_in;
        // the group's S2C, modraise and C2S run on one container of 2g polynomials
        const size_t N = c.N, Qin = in[i]->Ql();
        Ciphertext G;
        G.ctx = &c;
        G.resize(2 * g, in[i]->chain, s, false);  // every member's pair is copied in
        G.scale = d_in;
        G.ntt = true;
        for (size_t k = 0; k < g; k++)
            launch_copy_words(G.poly(2 * k), in[i + k]->data(), 2 * Qin * N, s);
        const Rotator rot = group_rotator(c, gk, s);
        slot_to_coeff(b, G, rot, s);
        ct_modraise(c, G, s);
        G.scale = q0;
        Ciphertext RE, IM;
        coeff_to_slot(b, G, RE, IM, rot, s);
        std::vector<Ciphertext> re(g), im(g);
        const size_t Qc = RE.Ql();
        for (size_t k = 0; k < g; k++) {
            for (int part = 0; part < 2; part++) {
                Ciphertext &dst = part ? im[k] : re[k];
                const Ciphertext &src = part ? IM : RE;
                dst.ctx = &c;
                dst.resize(2, src.chain, s, false);  // the copy writes both polynomials
                dst.scale = src.scale;
                dst.ntt = true;
                launch_copy_words(dst.data(), src.poly(2 * k), 2 * Qc * N, s);
            }
        }
        // amplitude correction (BatchBootstrap.cuh:444-460) folded into EvalMod:
        // sin(2 pi eps)/(2 pi) ~ eps, with the linear minimax constant over
        // |eps| <= E = F * scale_in * bound / q0 absorbing the cubic term (error
        // (2 pi)^2 E^3 / 24), times q0 / (F * scale_in) to return to the message.
        const double F = b.prm.layer3_factor;
        const double E = F * d_in * b.prm.input_bound / q0;
        const double pi = 3.14159265358979323846;
        const double cst = (1.0 + 0.5 * pi * pi * E * E) / (2.0 * pi) * q0 / (F * d_in);
        std::vector<Ciphertext *> parts;
        for (size_t k = 0; k < g; k++) {
            parts.push_back(&re[k]);
            parts.push_back(&im[k]);
        }
        eval_mod_batch(b, parts, rk, cst, osc, s);
        for (size_t k = 0; k < g; k++) {
            ct_mul_monomial_half(c, im[k], false, s);  // i * m_im
            add_checked(c, re[k], im[k], s);
            re[k].scale = osc;
            // copied into the caller's buffer (kept across calls, so a captured graph
            // writes the same memory on every replay)
            ct_clone(*out[i + k], re[k], s);
        }
        i = j;
    }
}

}  // namespace boot
}  // namespace shm
